"""Each benchmark configuration's physical set-up (turbulentbl_amd/cases.py:
Re and Ue from Re_theta,in, the 60 x 3 x 6 delta99 domain, tanh stretch 2.5,
the configuration's initial field and dt from CFL 0.3) on an oracle-sized
grid, through the C ABI against the oracle stepping the same state.

FP64 configurations run the oracle's orderings on both sides (z-line
red-black Gauss-Seidel, red-black SSOR-preconditioned GMRES): Gauss-Seidel
sweep counts must agree and the fields agree to the solver tolerance.  Mixed
configurations step the library's mixed path (FP32 fields inside FP64
refinement, fast preconditioner) against the FP64 oracle: velocity within
2e-6 and mean C_f within 1e-4 (the accuracy target of the mixed path)."""
import numpy as np
import pytest

from oracle import orc
from tests.case_util import gpu_from_case, make_config_case

pytestmark = pytest.mark.gpu

# oracle-sized grids (nX, nY span, nZ wall-normal) with each configuration's physics
GRIDS = {"c1": (64, 32, 48), "c2": (64, 16, 32), "c3": (64, 32, 48), "c4": (48, 16, 24), "c5a": (64, 32, 32),
         "c5b": (64, 24, 40)}
NSTEPS = 3


@pytest.fixture(scope="module")
def tbl():
    from turbulentbl_amd import tbl as t
    t.load()
    return t


@pytest.mark.parametrize("case", ["c1", "c2", "c4"])
def test_fp64_config_matches_oracle_same_orderings(tbl, case):
    g, s = make_config_case(case, GRIDS[case])
    print("%s: grid %s Re %.3f Ue %.4f dt %.4e nstep0 %d" % (case, GRIDS[case], g.Re, g.Ue, g.dt, s.nstep),
          flush=True)
    kw = dict(errorstop=1e-6, gs_mmax=1000, errtol=1e-9, kmax=30, mmax=1000)
    with gpu_from_case(g, s, kmax=30, errtol=1e-9, helm_mode=1, precond=tbl.TBL_PC_SSOR_RB) as sv:
        m = sv.mesh()
        dg = sv.step(NSTEPS)
        Ut, P, _ = sv.download()
        tau = sv.wall_shear()
    for k in ("x1", "x2", "x3", "Jac", "Jac2", "Ab", "At", "Ap"):
        assert np.array_equal(m[k], g.mesh[k]), k
    do = [s.step(gs_order=orc.GS_ZLINE_RB, pc=orc.PC_SSOR_RB, **kw) for _ in range(NSTEPS)]
    for a, b in zip(dg, do):
        assert a["gs_sweeps"] == [int(x) for x in b["gs_sweeps"]]
    assert np.abs(Ut - s.Ut).max() <= 1e-9 * np.abs(s.Ut).max()
    Pi = (slice(1, -1),) * 3
    assert np.abs(P[Pi] - s.Press[Pi]).max() <= 1e-7 * np.abs(s.Press[Pi]).max()
    tau_o = g.wall_shear(s.Ut)
    assert np.abs(tau - tau_o).max() <= 1e-8 * np.abs(tau_o).max()


@pytest.mark.parametrize("case", ["c3", "c5a", "c5b"])
def test_mixed_config_error_vs_fp64_oracle(tbl, case):
    g, s = make_config_case(case, GRIDS[case])
    print("%s: grid %s Re %.3f Ue %.4f dt %.4e" % (case, GRIDS[case], g.Re, g.Ue, g.dt), flush=True)
    kw = dict(errorstop=1e-6, gs_mmax=1000, errtol=1e-12, kmax=40, mmax=1000)
    with gpu_from_case(g, s, precision="mixed", kmax=40, errtol=1e-10, tol_mode=tbl.TBL_TOL_TRUE,
                       inner_tol=1e-4) as sv:
        dm = sv.step(NSTEPS)
        Ut, _, _ = sv.download()
        tau = sv.wall_shear()
    for _ in range(NSTEPS):
        s.step(gs_order=orc.GS_ZEBRA, pc=orc.PC_SSOR_RB, **kw)
    eu = np.abs(Ut - s.Ut).max() / np.abs(s.Ut).max()
    tau_o = g.wall_shear(s.Ut)
    cf = abs(tau[1:-1].mean() - tau_o[1:-1].mean()) / abs(tau_o[1:-1].mean())
    print("%s: u error %.2e, mean C_f error %.2e, true residuals %s" % (case, eu, cf,
                                                                       [d["gm_true_res"] for d in dm]))
    assert eu <= 2e-6
    assert cf <= 1e-4
