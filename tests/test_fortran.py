"""The Fortran iso_c_binding layer (turbulentbl_amd/fortran/tbl_gpu.f90) and
driver: struct layout against the C ABI (CPU), and the driver's time loop on
the GPU against the oracle stepping the same initial field."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "turbulentbl_amd", "fortran")
FC = "/opt/rocm/bin/amdflang"


def _have_fc():
    return os.path.exists(FC)


@pytest.mark.skipif(not _have_fc(), reason="amdflang not installed")
def test_fortran_params_layout_matches_c(tmp_path):
    from turbulentbl_amd import tbl
    if not os.path.exists(tbl.LIB_PATH):
        pytest.skip("library not built")
    b = str(tmp_path)
    subprocess.check_call([FC, "-O0", "-module-dir", b, "-c", os.path.join(FDIR, "tbl_gpu.f90"), "-o", b + "/m.o"])
    subprocess.check_call([FC, "-O0", "-I", b, "-module-dir", b, os.path.join(ROOT, "tests", "fortran", "layout_probe.f90"),
                           b + "/m.o", "-o", b + "/probe", "-L" + os.path.dirname(tbl.LIB_PATH), "-ltbl_hip",
                           "-Wl,-rpath," + os.path.dirname(tbl.LIB_PATH)])
    out = subprocess.check_output([b + "/probe"], text=True).split()
    p = tbl.default_params()
    import ctypes as C
    assert [int(x) for x in out[:3]] == [p.nX, p.nY, p.nZ]
    vals = [float(x) for x in out[3:10]]
    assert vals == [p.omega, p.Re, p.Ue, p.delta, p.errtol, p.errorstop, p.inner_tol]
    assert [int(x) for x in out[10:15]] == list(p.planes)
    assert [int(x) for x in out[15:24]] == [p.i_rec, p.freq_tim_avg, p.precision, p.gs_mmax, p.helm_mode,
                                           p.tol_mode, p.kmax, p.mmax, p.precond]
    assert int(out[24]) == C.sizeof(tbl.Params)


@pytest.mark.gpu
def test_fortran_driver_steps_match_oracle(tmp_path):
    """tbl_driver with the reference's initial field (dnssetbl.f90:151-166),
    block-by-block calls, 3 steps, against the oracle with the same orderings."""
    from oracle import orc
    drv = os.path.join(FDIR, "tbl_driver")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    nX, nY, nZ, nsteps = 40, 24, 30, 3
    out = tmp_path / "fields.bin"
    nml = tmp_path / "run.nml"
    nml.write_text("&tbl\n nX=%d, nY=%d, nZ=%d, nsteps=%d, i_rec=31, errtol=1e-11, errorstop=1e-13, "
                   "gs_mmax=400, prof_file='none', out_file='%s' /\n" % (nX, nY, nZ, nsteps, out))
    r = subprocess.run([drv, str(nml)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.strip() and not l.strip().startswith("checksum")]
    assert len(lines) == nsteps
    data = np.fromfile(out, dtype=np.float64)
    N = nX * nY * nZ
    Ut = data[:3 * N].reshape((nX, nY, nZ, 3), order="F")
    # the same initial field in the oracle
    planes = [max(2, round(5 * nX / 90.)), max(2, round(20 * nX / 90.)), max(2, round(40 * nX / 90.)),
              max(2, round(60 * nX / 90.)), min(nX - 1, round(85 * nX / 90.))]
    g = orc.Grid(nX, nY, nZ, i_rec=31, planes=planes)
    s = orc.State(g)
    x2, x3 = g.mesh["x2"], g.mesh["x3"]
    z = np.maximum(x3, 0.0)
    Ume = np.minimum(g.Ue, np.log(1 + 0.41 * z * g.Re) / 0.41 +
                     7.8 * (1 - np.exp(-z * g.Re / 11) - (z * g.Re / 11) * np.exp(-z * g.Re / 3)))
    pert = 0.01 * g.Ue * np.sin(10.0 * x2)[:, None] * np.cos(10.0 * x3)[None, :]
    s.Ut[..., 0] = Ume[None, None, :] + pert[None]
    s.Ut[..., 1] = pert[None]
    s.Ut[..., 2] = pert[None]
    s.Ume[:] = Ume
    for _ in range(nsteps):
        d = s.step(gs_order=orc.GS_ZLINE_RB, pc=orc.PC_SSOR_RB, errorstop=1e-13, gs_mmax=400, errtol=1e-11)
    assert np.abs(Ut - s.Ut).max() <= 1e-8 * np.abs(s.Ut).max()
    div = float(lines[-1].split()[2])
    assert div == pytest.approx(d["div"], rel=1e-5, abs=1e-6)
