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
    assert [int(x) for x in out[25:27]] == [tbl.Params.scheme.offset, tbl.Params.forc.offset]
    assert int(out[27]) == p.scheme and [float(x) for x in out[28:31]] == list(p.forc)
    assert [int(x) for x in out[31:33]] == [tbl.Params.slabs.offset, p.slabs]


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


def _restart_run(tmp_path, nsteps, freq=500, extra=""):
    """Reference input files (Input/3D.dat, statData.dat, initProf.dat) of a
    seeded 40x24x30 case at nstep 1000, the driver restarted from them
    (prmpt = 1), and the oracle stepped from exactly what the files hold."""
    from oracle import orc
    from tests.case_util import make_case
    from turbulentbl_amd import refio
    drv = os.path.join(FDIR, "tbl_driver")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    nX, nY, nZ, i_rec = 40, 24, 30, 31
    g, s = make_case((nX, nY, nZ), freq=freq, seed=5, nstep0=1000)
    s.nstep = 1000
    inp = tmp_path / "Input"
    inp.mkdir()
    m = g.mesh
    refio.write_3d(inp / "3D.dat", m["x1"], m["x2"], m["x3"], s.Ut, s.Press, 0.1, 1000)
    rec = i_rec - 1
    refio.write_stat(inp / "statData.dat", s.avg[rec, :, 0], s.avg[rec, :, 1], s.avg[rec, :, 2], s.Wt[0, :], s.newk)
    nprof = min(37, nZ)
    refio.write_init_prof(inp / "initProf.dat", s.Ume[:nprof])
    dump = tmp_path / "out3D.dat"
    nml = tmp_path / "run.nml"
    nml.write_text("&tbl\n nX=%d, nY=%d, nZ=%d, nsteps=%d, i_rec=%d, errtol=1e-11, errorstop=1e-13, gs_mmax=400, "
                   "freqTimAvg=%d, prmpt=1, nprof=%d, prof_file='%s', in_file='%s', stat_file='%s', dump_file='%s' %s/\n"
                   % (nX, nY, nZ, nsteps, i_rec, freq, nprof, inp / "initProf.dat", inp / "3D.dat",
                      inp / "statData.dat", dump, extra))
    r = subprocess.run([drv, str(nml)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    t0, n0, _, _, _, U0, P0 = refio.read_3d(inp / "3D.dat", nX, nY, nZ)
    u, v, w, wt, newk = refio.read_stat(inp / "statData.dat", nZ)
    o = orc.State(g)
    o.Ut[...] = U0
    o.Press[...] = P0
    o.avg[rec, :, 0], o.avg[rec, :, 1], o.avg[rec, :, 2] = u, v, w
    o.Wt[0, :] = wt
    o.newk[:] = newk
    o.Ume[:nprof] = refio.read_init_prof(inp / "initProf.dat", nprof)
    o.Ume[nprof:] = o.Ume[nprof - 1]
    for p in range(3):
        o.Ufluc_rec[1:-1, 1:-1, p] = U0[rec, 1:-1, 1:-1, p] - o.avg[rec, None, 1:-1, p]
    o.nstep, o.time = n0, t0
    for _ in range(nsteps):
        o.step(gs_order=orc.GS_ZLINE_RB, pc=orc.PC_SSOR_RB, errorstop=1e-13, gs_mmax=400, errtol=1e-11)
    return g, o, dump


@pytest.mark.gpu
def test_fortran_driver_restart_from_reference_files(tmp_path):
    """prmpt = 1: the driver reads Input/3D.dat and Input/statData.dat in the
    reference's formats (dnssetbl.f90:118-150), steps past the averaging
    window (recycling active), and dumps the combined restart format; the
    oracle steps the same state read back from the same files."""
    from turbulentbl_amd import refio
    nsteps = 2
    g, o, dump = _restart_run(tmp_path, nsteps)
    t1, n1, _, _, _, U1, P1 = refio.read_3d(dump, g.nX, g.nY, g.nZ)
    assert n1 == 1000 + nsteps
    assert np.abs(U1 - o.Ut).max() <= 1e-8 * np.abs(o.Ut).max()


@pytest.mark.gpu
def test_fortran_driver_writes_rescale_statistics(tmp_path):
    """A statistics rescale inside the run (freqTimAvg = 2, steps 1001-1002):
    the driver's Output files (dnssetbl.f90:1071-1198) hold the oracle's
    averages, Reynolds stresses, Lund weight, new(k), eta(k), lamda and the
    integral thicknesses, line for line in the reference's formats (the
    same files written by refio from the oracle state)."""
    from turbulentbl_amd import refio
    outd = tmp_path / "Output"
    outd.mkdir()
    g, o, _ = _restart_run(tmp_path, 2, freq=2, extra="out_dir='%s' " % outd)
    expd = tmp_path / "expected"
    expd.mkdir()
    m = g.mesh
    refio.write_rescale_outputs(str(expd), m["x1"], m["x3"], o.avg, o.Wt, o.newk, o.bltkns, o.deltastar, o.theta,
                                o.lamda, o.lamda03, o.time, g.Ue, g.Re, g.i_rec,
                                statdata=str(expd / "statData.dat"))
    for name in ("lamda.dat", "W.dat", "boundLyrThik.dat", "newK.dat", "RMS.dat", "meanVel.dat", "statData.dat"):
        got = (outd / name).read_text().splitlines()
        exp = (expd / name).read_text().splitlines()
        assert len(got) == len(exp), name
        assert [len(l) for l in got] == [len(l) for l in exp], name          # field widths
        hdr = name not in ("lamda.dat", "statData.dat")
        if hdr:
            assert got[0] == exp[0], name
        if name == "statData.dat":
            a = np.array([[float(l[1 + 16 * q:16 + 16 * q]) for q in range(4)] + [int(l[64:69])] for l in got])
            b = np.array([[float(l[1 + 16 * q:16 + 16 * q]) for q in range(4)] + [int(l[64:69])] for l in exp])
        else:
            a, b = refio.read_table(outd / name, hdr), refio.read_table(expd / name, hdr)
        if name == "newK.dat":
            assert np.array_equal(a[:, :2], b[:, :2])
        # printed values agree to the last printed digit (F15.7 / F20.10)
        assert np.abs(a - b).max() <= 2e-7 * max(1.0, np.abs(b).max()), name


@pytest.mark.gpu
@pytest.mark.parametrize("blocks", [1, 0])
def test_fortran_driver_resumes_from_a_checkpoint_bitwise(tmp_path, blocks):
    """tbl_driver run straight for 5 steps equals 2 steps + ckpt_out, then a
    second run with ckpt_in for 3 steps, bit for bit (freqTimAvg = 3: the
    run crosses a rescale and the start of the recycled inflow and of the KE
    guard), both with block-by-block calls and with whole steps."""
    drv = os.path.join(FDIR, "tbl_driver")
    if not os.path.exists(drv):
        pytest.skip("driver not built")
    base = ("nX=40, nY=24, nZ=30, i_rec=31, freqTimAvg=3, errtol=1e-10, prof_file='none', blocks=%d, "
            "tau_window=2, " % blocks)

    def run(name, extra):
        nml = tmp_path / (name + ".nml")
        nml.write_text("&tbl\n " + base + extra + " /\n")
        r = subprocess.run([drv, str(nml)], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
        assert r.returncode == 0, r.stdout + r.stderr
        return r.stdout

    full = tmp_path / "full.bin"
    part = tmp_path / "part.bin"
    ck = tmp_path / "run.ckpt"
    out_full = run("full", "nsteps=5, out_file='%s'" % full)
    run("first", "nsteps=2, ckpt_out='%s'" % ck)
    out_rest = run("rest", "nsteps=3, ckpt_in='%s', out_file='%s'" % (ck, part))
    assert full.read_bytes() == part.read_bytes()
    # the printed time / step / divergence lines of steps 3..5 are the same
    assert out_full.strip().splitlines()[-3:] == out_rest.strip().splitlines()[-3:]
