"""GPU discrete adjoint (tbl_adjoint_window) and the direct Helmholtz solve.

* TBL_HELM_DIRECT equals the oracle's Gauss-Seidel converged to machine
  precision.
* The window's forward map equals the FP64 CPU exact-solve step
  (tests/adjoint_ref.py).
* The adjoint gradient equals the FP64 CPU reference adjoint.
* The gradient does not depend on the checkpoint count.
* Through the torch.autograd harness, the gradient matches central finite
  differences along 8 seeded directions, over a 20-step window and over the
  full 200-step window with 20 checkpoints.
"""
import numpy as np
import pytest

from oracle import orc
from tests.adjoint_ref import ExactStep, forward, objective, reference_adjoint
from tests.case_util import C1_GRID, RAGGED, REF_GRID, gpu_from_case, load_case, make_case

pytestmark = pytest.mark.gpu

TINY = (12, 8, 10)


@pytest.fixture(scope="module")
def tbl():
    from turbulentbl_amd import tbl as t
    t.load()
    return t


def _vw(g, seed, scale=1e-3):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(scale * g.Ue * rng.standard_normal((g.nX, g.nY)))


@pytest.mark.parametrize("grid", [RAGGED, C1_GRID, REF_GRID])
def test_direct_helmholtz_matches_converged_gauss_seidel(tbl, grid):
    g, s = make_case(grid, seed=2)
    _, Q = g.momentum_rhs(s.Ut, s.Press)
    with gpu_from_case(g, s, helm_mode=tbl.TBL_HELM_DIRECT) as sv:
        sv.set_field(tbl.F_Q, Q)
        sw, _ = sv.diffusion_solve(1e-6, 1000)
        Ut = sv.get_field(tbl.F_UT)
    Uo = s.Ut.copy(order="F")
    g.diffusion_solve(Q, Uo, order=orc.GS_ZLINE_RB, errorstop=1e-15, mmax=400)
    assert list(sw) == [2, 2, 2]
    assert np.abs(Ut - Uo).max() <= 1e-12 * np.abs(Uo).max()


def test_window_forward_is_the_exact_step(tbl):
    g, s = make_case(TINY, seed=3)
    vw = _vw(g, 3)
    n, i0, i1 = 3, 5, 11
    step = ExactStep(g, s)
    states = forward(step, s.Ut, s.Press, vw, s.nstep, n)
    Jc = objective(g, states, i0, i1)
    with gpu_from_case(g, s) as sv:
        J, gr, nf, na = sv.adjoint_window(n, i0, i1, wall_w=vw, forward_only=True)
        Ut, P, _ = sv.download()
        assert sv.get_state()["nstep"] == s.nstep + n
    assert gr is None and nf == n and na == 0
    assert abs(J - Jc) <= 1e-11 * abs(Jc)
    assert np.abs(Ut - states[-1][0]).max() <= 1e-11 * np.abs(states[-1][0]).max()
    assert np.abs(P - states[-1][1]).max() <= 1e-9 * np.abs(states[-1][1]).max()


def test_adjoint_matches_cpu_reference_adjoint(tbl):
    g, s = make_case(TINY, seed=4)
    vw = _vw(g, 4)
    n, i0, i1 = 2, 5, 11
    Jr, gref = reference_adjoint(ExactStep(g, s), s.Ut, s.Press, vw.copy(order="F"), s.nstep, n, i0, i1)
    with gpu_from_case(g, s) as sv:
        J, grad, nf, na = sv.adjoint_window(n, i0, i1, wall_w=vw, nsnap=1)
    assert abs(J - Jr) <= 1e-11 * abs(Jr)
    assert na == n
    scale = np.abs(gref).max()
    err = np.abs(grad - gref).max() / scale
    print("max |g_gpu - g_ref| / max |g_ref| = %.3e" % err)
    assert err <= 1e-6


def test_gradient_independent_of_checkpoints(tbl):
    g, s = make_case(C1_GRID, seed=5)
    vw = _vw(g, 5)
    n, i0, i1 = 7, 30, 58
    out = []
    for nsnap in (1, 2, 6):
        with gpu_from_case(g, s) as sv:
            out.append(sv.adjoint_window(n, i0, i1, wall_w=vw, nsnap=nsnap))
    for J, gr, nf, na in out[1:]:
        assert J == out[0][0]
        assert np.array_equal(gr, out[0][1])
        assert na == n
    # fewer snapshots cost more recomputation: n + the binomial schedule's re-advances
    nfs = [o[2] for o in out]
    assert nfs[0] >= nfs[1] >= nfs[2] >= n


def test_autograd_gradient_matches_central_differences(tbl):
    """The 256x96x64 case (nX=256, nY=64 span, nZ=96 wall-normal), 20 steps:
    8 seeded directions, central differences with eps = 1e-6 in velocity
    units (directions scaled by U_e), evaluated with the base trajectory's
    Peclet-switch pattern replayed (the switch is a jump of the reference's
    scheme; a difference across it is not a derivative)"""
    import torch
    from turbulentbl_amd.autograd import wall_shear_objective
    g, s = make_case((256, 64, 96), seed=6)
    nX, nY = g.nX, g.nY
    n, i0, i1 = 20, int(0.6 * (nX - 1)) + 1, int(0.9 * (nX - 1)) + 1
    # control on x in [0.2, 0.4] Lx, interior span
    x = g.mesh["x1"]
    ctrl = np.zeros((nX, nY))
    ctrl[(x >= 0.2 * x[-1]) & (x <= 0.4 * x[-1]), 1:nY - 1] = 1.0
    mask = torch.from_numpy(ctrl)
    base = torch.from_numpy(_vw(g, 6)) * mask
    with gpu_from_case(g, s) as sv:
        def reset(solver):
            load_case(solver, s)

        w = base.clone().requires_grad_(True)
        J = wall_shear_objective(w, sv, reset, n, i0, i1, nsnap=4, freeze=True)
        J.backward()
        grad = w.grad.clone()
        rng = np.random.default_rng(8)
        eps = 1e-6
        errs, free = [], []
        for q in range(8):
            d = torch.from_numpy(rng.standard_normal((nX, nY))) * mask * g.Ue
            ad = float((grad * d).sum())
            with torch.no_grad():
                Jp = wall_shear_objective(base + eps * d, sv, reset, n, i0, i1, freeze=True)
                Jm = wall_shear_objective(base - eps * d, sv, reset, n, i0, i1, freeze=True)
                errs.append(abs(float((Jp - Jm) / (2 * eps)) - ad) / abs(ad))
                Jp = wall_shear_objective(base + eps * d, sv, reset, n, i0, i1)
                Jm = wall_shear_objective(base - eps * d, sv, reset, n, i0, i1)
                free.append(abs(float((Jp - Jm) / (2 * eps)) - ad) / abs(ad))
    print("central differences, switches replayed: ", ["%.2e" % e for e in errs])
    print("central differences, switches evaluated:", ["%.2e" % e for e in free])
    assert max(errs) <= 1e-5


def test_autograd_gradient_full_window(tbl):
    """The 256x96x64 gradient check at its full window: 200 steps with 20
    binomial checkpoints, 8 seeded directions, central differences with
    eps = 1e-6 (Peclet switches replayed, as above)"""
    import torch
    from turbulentbl_amd.autograd import wall_shear_objective
    g, s = make_case((256, 64, 96), seed=6)
    nX, nY = g.nX, g.nY
    n, i0, i1 = 200, int(0.6 * (nX - 1)) + 1, int(0.9 * (nX - 1)) + 1
    x = g.mesh["x1"]
    ctrl = np.zeros((nX, nY))
    ctrl[(x >= 0.2 * x[-1]) & (x <= 0.4 * x[-1]), 1:nY - 1] = 1.0
    mask = torch.from_numpy(ctrl)
    base = torch.from_numpy(_vw(g, 6)) * mask
    with gpu_from_case(g, s) as sv:
        def reset(solver):
            load_case(solver, s)

        w = base.clone().requires_grad_(True)
        J = wall_shear_objective(w, sv, reset, n, i0, i1, nsnap=20, freeze=True)
        J.backward()
        grad = w.grad.clone()
        rng = np.random.default_rng(8)
        eps = 1e-6
        errs = []
        for q in range(8):
            d = torch.from_numpy(rng.standard_normal((nX, nY))) * mask * g.Ue
            ad = float((grad * d).sum())
            with torch.no_grad():
                Jp = wall_shear_objective(base + eps * d, sv, reset, n, i0, i1, freeze=True)
                Jm = wall_shear_objective(base - eps * d, sv, reset, n, i0, i1, freeze=True)
            errs.append(abs(float((Jp - Jm) / (2 * eps)) - ad) / abs(ad))
    print("200 steps, central differences, switches replayed:", ["%.2e" % e for e in errs])
    assert max(errs) <= 1e-5


# ---------------------------------------------------------------- mixed forward, FP64-accumulated adjoint
@pytest.mark.parametrize("grid,n,nsnap", [(TINY, 4, 2), ((48, 16, 24), 12, 3)])
def test_mixed_window_adjoint_against_fp64(tbl, grid, n, nsnap):
    """A mixed context's window: forward in mixed precision (FP32 fields and
    fast solves, pressure refined in FP64), FP32 checkpoints, adjoint
    accumulated in FP64 about the mixed trajectory.  Against the FP64
    context's window on the same state and control: J within 1e-6 and the
    gradient within 1e-4 relative (max norm; the two linearise about
    trajectories that differ by FP32 rounding).  The gradient does not depend
    on the checkpoint count (the recomputed mixed forward is deterministic)."""
    g, s = make_case(grid, seed=5)
    vw = _vw(g, 5)
    i0, i1 = int(0.6 * (g.nX - 1)) + 1, int(0.9 * (g.nX - 1)) + 1
    with gpu_from_case(g, s) as sv:
        J64, g64, _, _ = sv.adjoint_window(n, i0, i1, wall_w=vw, nsnap=nsnap)
    out = []
    for ns in (nsnap, n):
        with gpu_from_case(g, s, precision="mixed") as sv:
            out.append(sv.adjoint_window(n, i0, i1, wall_w=vw, nsnap=ns))
    (Jm, gm, nf, na), (Jm2, gm2, _, _) = out
    ej = abs(Jm - J64) / abs(J64)
    eg = np.abs(gm - g64).max() / np.abs(g64).max()
    print("mixed vs FP64 window %s, %d steps: J %.3e, gradient %.3e (forward steps %d)" % (grid, n, ej, eg, nf))
    assert ej <= 1e-6
    assert eg <= 1e-4
    assert na == n and nf >= n
    assert np.array_equal(gm, gm2) and Jm == Jm2


def test_mixed_window_adjoint_against_cpu_reference(tbl):
    """The mixed window's FP64-accumulated adjoint against the CPU FP64
    reference adjoint (dense exact solves) on the tiny grid, in the case the
    FP64 window matches it to 1e-6 (test_adjoint_matches_cpu_reference_adjoint;
    one step more puts a Peclet switch inside the reference's +-eps
    differences): within 1e-4."""
    g, s = make_case(TINY, seed=4)
    vw = _vw(g, 4)
    n, i0, i1 = 2, 5, 11
    Jr, gref = reference_adjoint(ExactStep(g, s), s.Ut, s.Press, vw.copy(order="F"), s.nstep, n, i0, i1)
    with gpu_from_case(g, s, precision="mixed") as sv:
        J, grad, _, _ = sv.adjoint_window(n, i0, i1, wall_w=vw, nsnap=1)
    eg = np.abs(grad - gref).max() / np.abs(gref).max()
    print("mixed window vs CPU reference adjoint: J %.3e, gradient %.3e" % (abs(J - Jr) / abs(Jr), eg))
    assert abs(J - Jr) <= 1e-6 * abs(Jr)
    assert eg <= 1e-4
