"""The reference's file formats (turbulentbl_amd/refio.py): byte equality
with what Fortran's edit descriptors write (amdflang on
tests/fortran/format_probe.f90), and write/read round trips."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from turbulentbl_amd import refio as r

HERE = os.path.dirname(os.path.abspath(__file__))
FC = os.environ.get("TBL_FC", "/opt/rocm/bin/amdflang")


def _python_lines():
    out = [r.record104(0.5, 12.0, 0.0013, 19.798123456789012, -0.000123456789, 0.0, 1.5e-120),
           r.record104(11.9882697947214, 3.96825396825397, 3.4, -1.0e-300, 2.5e300, 1.0, -9.99999999999999999e-1),
           'TITLE ="' + r._F(1.2345, 20, 10) + r._I(1500, 9) + '"',
           'TITLE ="' + r._F(1.2345, 20, 10) + r._I(1500, 9) + r._I(1, 3) + '"',
           "ZONE k=" + r._I(1024, 4) + ",j=" + r._I(256, 4) + ",i=" + r._I(192, 4) + ",DATAPACKING=POINT",
           "ZONE k=" + r._I(1024, 4) + ",j=" + r._I(254, 4) + ",i=" + r._I(192, 4) + ',DATAPACKING="POINT"',
           "".join(" " + r._F(a, 15, 7) for a in (19.5, -0.25, 0.125, 0.99)) + r._I(7, 5, 3),
           "".join(" " + r._F(a, 15, 7) for a in (0.0, 1.0e-8, -19.798, 1.0)) + r._I(123, 5, 3),
           " " * 24 + r._E(19.798, 18, 8),
           "".join("  " + r._F(a, 20, 10) for a in (0.1, 1.0000000001, -2.0)),
           "".join(" " + r._F(a, 15, 7) for a in (-1.0e-8, float("nan"), float("inf"))),
           "".join(" " + r._F(a, 15, 7) for a in (-float("inf"), 1e9, -1.5)),
           "".join("  " + r._F(a, 20, 10) for a in (1.0, 0.0123456789012, 2.5, 3.0, 4.0, 500.25, 0.0)),
           r._I(7, 5) + r._I(12, 5) + r._F(1.25, 20, 10),
           "".join(" " + r._F(a, 15, 7) for a in (1.0, 2.0, 3.0, 4.0, 5.0)),
           "".join(" " + r._F(a, 15, 7) for a in (1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, -8.0)),
           r._list_header('Variables = "x1","x3","Wt"').rstrip("\n")]
    return out


@pytest.mark.skipif(not (os.path.exists(FC) or shutil.which(FC)), reason="no Fortran compiler")
def test_writers_match_fortran_edit_descriptors(tmp_path):
    exe = tmp_path / "probe"
    subprocess.check_call([FC, os.path.join(HERE, "fortran", "format_probe.f90"), "-o", str(exe)], cwd=tmp_path)
    got = subprocess.check_output([str(exe)], cwd=tmp_path).decode().splitlines()
    assert got == _python_lines()


def test_3d_roundtrip_and_layout(tmp_path):
    rng = np.random.default_rng(0)
    nX, nY, nZ = 5, 6, 4
    x1, x2, x3 = np.linspace(0, 12, nX), np.linspace(0, 4, nY), np.linspace(0, 3.4, nZ)
    Ut = np.asfortranarray(rng.standard_normal((nX, nY, nZ, 3)) * 20)
    P = np.asfortranarray(rng.standard_normal((nX, nY, nZ)))
    path = tmp_path / "3D.dat"
    r.write_3d(path, x1, x2, x3, Ut, P, 1.25, 1500)
    lines = open(path).read().split("\n")
    # header, variables, zone; then per i a blank line, per j a blank line, nZ records
    assert lines[0] == 'TITLE ="        1.2500000000     1500"'
    assert lines[2].startswith("ZONE k=   5,j=   6,i=   4")
    assert lines[3] == "" and lines[4] == "" and len(lines[5]) == 3 * 11 + 4 * 23
    t, n, a1, a2, a3, U2, P2 = r.read_3d(path, nX, nY, nZ)
    assert (t, n) == (1.25, 1500)
    assert np.allclose(U2, Ut, rtol=1e-14, atol=0) and np.allclose(P2, P, rtol=1e-14, atol=0)
    assert np.allclose(a1, x1, atol=5e-7) and np.allclose(a3, x3, atol=5e-7)


def test_rank_dump_layout(tmp_path):
    nX, nY, nZ = 3, 6, 4
    z = np.zeros((nX, nY, nZ))
    r.write_3d_rank(tmp_path / "3D001.dat", np.arange(nX), np.arange(nY), np.arange(nZ),
                    np.zeros((nX, nY, nZ, 3)), z, 0.5, 10)
    lines = open(tmp_path / "3D001.dat").read().split("\n")
    assert lines[0] == 'TITLE ="        0.5000000000       10  1"'
    assert lines[2] == 'ZONE k=   3,j=   4,i=   4,DATAPACKING="POINT"'
    assert sum(1 for ln in lines if len(ln) > 80) == nX * (nY - 2) * nZ


def test_stat_and_profile_roundtrip(tmp_path):
    nZ = 7
    rng = np.random.default_rng(1)
    u, v, w, wt = (rng.uniform(-20, 20, nZ) for _ in range(4))
    newk = np.arange(1, nZ + 1, dtype=np.int32)
    r.write_stat(tmp_path / "statData.dat", u, v, w, wt, newk)
    u2, v2, w2, wt2, nk2 = r.read_stat(tmp_path / "statData.dat", nZ)
    for a, b in ((u, u2), (v, v2), (w, w2), (wt, wt2)):
        assert np.allclose(a, b, atol=5e-8, rtol=0)
    assert np.array_equal(nk2, newk)
    prof = rng.uniform(0, 20, 37)
    r.write_init_prof(tmp_path / "initProf.dat", prof)
    assert np.allclose(r.read_init_prof(tmp_path / "initProf.dat"), prof, rtol=6e-8, atol=0)   # 8 digits


def test_rescale_outputs_layout(tmp_path):
    """the files of one statistics rescale: headers, record counts, columns,
    Reynolds stresses from the nine averages, eta from the Lund weight at the
    recycling plane, lamda.dat appended"""
    rng = np.random.default_rng(3)
    nX, nZ, i_rec = 6, 5, 4
    x1, x3 = np.linspace(0, 12, nX), np.linspace(0, 3.4, nZ)
    avg = rng.uniform(0.1, 1.0, (nX, nZ, 9))
    Wt = rng.uniform(0.5, 1.2, (nX, nZ))
    newk = np.array([1, 2, 4, 4, 4], dtype=np.int32)
    bl, ds, th = rng.uniform(1, 2, nX), rng.uniform(0.1, 0.2, nX), rng.uniform(0.05, 0.1, nX)
    for t in (0.5, 1.0):
        r.write_rescale_outputs(str(tmp_path), x1, x3, avg, Wt, newk, bl, ds, th, 1.1, 0.9, t, 19.798, 150.92, i_rec,
                                statdata=str(tmp_path / "statData.dat"))
    lam = r.read_table(tmp_path / "lamda.dat", header=False)
    assert lam.shape == (2, 3) and list(lam[:, 0]) == [0.5, 1.0]
    rms = r.read_table(tmp_path / "RMS.dat")
    assert open(tmp_path / "RMS.dat").readline() == ' Variables = "x1","x3","Ruu","Rvv","Rww","Ruv","Ruw","Rvw"\n'
    assert rms.shape == (nX * nZ, 8)
    ruv = avg[..., 6] - avg[..., 0] * avg[..., 1]
    assert np.abs(rms[:, 5] - ruv.reshape(-1)).max() <= 5e-8          # i outer, k inner
    nk = r.read_table(tmp_path / "newK.dat")
    assert list(nk[:, 0]) == [1, 2, 3, 4, 5] and list(nk[:, 1]) == list(newk)
    eta = np.where(Wt[i_rec - 1, :nZ - 1] <= 0.99, 1.1, 0.9)
    assert np.allclose(nk[:nZ - 1, 2], eta) and nk[-1, 2] == 0.0
    blt = r.read_table(tmp_path / "boundLyrThik.dat")
    assert np.allclose(blt[:, 4], ds / th, atol=1e-9) and np.all(blt[:, 6] == 0.0)
    u, v, w, wt, k = r.read_stat(tmp_path / "statData.dat", nZ)
    assert np.allclose(u, avg[i_rec - 1, :, 0], atol=5e-8) and list(k) == list(newk)
