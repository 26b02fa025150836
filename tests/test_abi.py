"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/tbl.h declares (no compute: there is no GPU here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "tbl.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tbl_[a-z0-9_]+)\s*\(", src)))


def _lib():
    from turbulentbl_amd import tbl
    if not os.path.exists(tbl.LIB_PATH):
        pytest.skip("libtbl_hip.so not built")
    return tbl.load()


def test_header_declares_the_expected_surface():
    names = _declared()
    for n in ("tbl_create", "tbl_momentum_rhs", "tbl_ppe_solve", "tbl_project", "tbl_recycle_inflow",
              "tbl_step", "tbl_wall_shear"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib()
    for n in _declared():
        assert hasattr(lib, n), n


def test_binding_lists_every_declared_symbol():
    from turbulentbl_amd import tbl
    assert sorted(tbl.EXPORTS) == _declared()


def _header_fields(src, name):
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = decl.split(None, 1)[1]
        for nm in names.split(","):
            fields.append(re.sub(r"\[.*\]", "", nm).strip().lstrip("*"))
    return fields


@pytest.mark.parametrize("cname,pyname", [("tbl_params", "Params"), ("tbl_state", "State"),
                                          ("tbl_diag", "Diag"), ("tbl_adjoint_opts", "AdjointOpts")])
def test_struct_layouts_match_header(cname, pyname):
    from turbulentbl_amd import tbl
    src = open(os.path.join(ROOT, "include", "tbl.h")).read()
    assert [f[0] for f in getattr(tbl, pyname)._fields_] == _header_fields(src, cname)


def test_default_params_are_the_reference_parameters():
    lib = _lib()
    from turbulentbl_amd import tbl
    p = tbl.Params()
    lib.tbl_default_params(C.byref(p))
    assert (p.nX, p.nY, p.nZ) == (90, 80, 50)
    assert (p.dt, p.Re, p.Ue, p.delta) == (1.0e-4, 150.92, 19.798, 1.95)
    assert p.omega == 1.83500003814697265625            # REAL(4) literal
    assert (p.kmax, p.mmax, p.errtol, p.errorstop) == (50, 1000, 1.0e-5, 1.0e-6)
    assert p.i_rec == 68 and p.freq_tim_avg == 500 and list(p.planes) == [5, 20, 40, 60, 85]


INVALID = [(dict(nY=81), "nY must be even"), (dict(kmax=65), "kmax"), (dict(nX=60), "i_rec"),
           (dict(planes=(5, 20, 40, 60, 95)), "mass-flux planes"), (dict(dt=0.0), "dt and Re"),
           (dict(Re=-1.0), "dt and Re"), (dict(tol_mode=7), "tol_mode"), (dict(precision=5), "precision"),
           (dict(nZ=4), "at least 8")]


@pytest.mark.parametrize("bad,msg", INVALID)
def test_create_rejects_invalid_params_before_touching_a_device(bad, msg):
    lib = _lib()
    from turbulentbl_amd import tbl
    p = tbl.default_params(**bad)
    h = C.c_void_p()
    assert lib.tbl_create(C.byref(p), C.byref(h)) == -1
    assert not h.value
    # no context exists: the message is kept for tbl_last_error(NULL, ...)
    buf = C.create_string_buffer(256)
    assert lib.tbl_last_error(None, buf, 256) == 0
    assert msg in buf.value.decode()
    with pytest.raises(tbl.TblError, match=msg):
        sizes = dict(nX=p.nX, nY=p.nY, nZ=p.nZ)
        tbl.Solver(**{**sizes, **bad})
