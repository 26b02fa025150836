"""The FP64 CPU reference adjoint (tests/adjoint_ref.py) against the oracle's
own step and against central differences of its forward map."""
import numpy as np

from oracle import orc
from tests.adjoint_ref import ExactStep, forward, objective, reference_adjoint
from tests.case_util import make_case

TINY = (12, 8, 10)


def _case(seed=3):
    g, s = make_case(TINY, seed=seed)
    rng = np.random.default_rng(seed)
    vw = np.asfortranarray(1e-3 * g.Ue * rng.standard_normal((g.nX, g.nY)))
    return g, s, vw


def test_exact_step_is_the_oracle_step_at_tight_tolerance():
    g, s, vw = _case()
    step = ExactStep(g, s)
    U1, P1 = step(s.Ut, s.Press, vw, s.nstep + 1)
    s.wall_w = vw
    s.step(gs_order=orc.GS_ZLINE_RB, errorstop=1e-14, gs_mmax=3000, pc=orc.PC_SSOR_RB, errtol=1e-14, kmax=50,
           mmax=200)
    assert np.abs(U1 - s.Ut).max() <= 1e-10 * np.abs(s.Ut).max()
    assert np.abs(P1 - s.Press).max() <= 1e-8 * np.abs(s.Press).max()


def test_reference_adjoint_matches_central_differences():
    g, s, vw = _case(4)
    step = ExactStep(g, s)
    n, i0, i1 = 2, 5, 11
    J, grad = reference_adjoint(step, s.Ut, s.Press, vw.copy(order="F"), s.nstep, n, i0, i1)
    rng = np.random.default_rng(9)
    for _ in range(2):
        d = np.asfortranarray(rng.standard_normal(vw.shape))
        eps = 1e-3
        Jp = objective(g, forward(step, s.Ut, s.Press, vw + eps * d, s.nstep, n), i0, i1)
        Jm = objective(g, forward(step, s.Ut, s.Press, vw - eps * d, s.nstep, n), i0, i1)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.sum(grad * d))
        assert abs(fd - ad) <= 1e-6 * abs(ad), (fd, ad)
    # J itself
    assert np.isclose(J, objective(g, forward(step, s.Ut, s.Press, vw, s.nstep, n), i0, i1), rtol=1e-14)
