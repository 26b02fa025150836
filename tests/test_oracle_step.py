"""The oracle's composed time step: determinism and independence of the
solver orderings / slab count at tight tolerance (the converged solution of
each linear solve is unique, so orderings may only differ by the tolerance)."""
import numpy as np
import pytest

from oracle import orc
from turbulentbl_amd import inflow


def make_state(nX=24, nY=12, nZ=16, P=1, freq=500, seed=1, nstep0=1000):
    i_rec = int(round(0.8 * (nX - 1)))
    planes = tuple(max(2, min(nX - 1, int(round(p / 90.0 * nX)))) for p in orc.REF_PLANES)
    g = orc.Grid(nX, nY, nZ, P=P, i_rec=i_rec, planes=planes, freq_tim_avg=freq)
    s = orc.State(g)
    s.Ut[...] = inflow.initial_field(g.mesh, g.Re, g.Ue, seed=seed)
    rs = inflow.recycle_state(g.mesh["x3"], nX, g.Re, g.Ue, i_rec)
    s.avg[...] = rs["avg"]
    s.Wt[...] = rs["Wt"]
    s.newk[...] = rs["newk"]
    s.Ume[...] = rs["Ume"]
    s.bltkns[...] = rs["bltkns"]
    for p in range(3):
        s.Ufluc_rec[:, :, p] = s.Ut[i_rec - 1, :, :, p] - s.avg[i_rec - 1, None, :, p]
    s.nstep = nstep0
    return g, s


TIGHT = dict(errorstop=1e-13, errtol=1e-13, kmax=30)


def test_step_is_deterministic():
    out = []
    for _ in range(2):
        g, s = make_state()
        for _ in range(3):
            s.step()
        out.append((s.Ut.copy(), s.Press.copy(), s.sums.copy()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", [dict(P=2), dict(gs_order=orc.GS_RB, pc=orc.PC_SSOR_RB),
                                     dict(pc=orc.PC_JACOBI)])
def test_step_independent_of_ordering_and_slabs_at_tight_tolerance(variant):
    P = variant.pop("P", 1)
    g1, s1 = make_state(P=1)
    g2, s2 = make_state(P=P)
    for _ in range(2):
        d1 = s1.step(**TIGHT)
        d2 = s2.step(**TIGHT, **variant)
    u1, u2 = s1.Ut, s2.Ut
    assert np.abs(u1 - u2).max() <= 1e-9 * np.abs(u1).max()
    p1, p2 = s1.Press[1:-1, 1:-1, 1:-1], s2.Press[1:-1, 1:-1, 1:-1]
    assert np.abs(p1 - p2).max() <= 1e-8 * np.abs(p1).max()
    assert d1["div"] == pytest.approx(d2["div"], rel=1e-6, abs=1e-10)


def test_rescale_and_ke_path():
    g, s = make_state(freq=3, nstep0=0)
    s.cnt = 0
    seen = []
    for _ in range(5):
        d = s.step()
        seen.append(d["rescaled"])
        assert np.isfinite(s.Ut).all()
    assert seen == [False, False, True, False, False]
    # after the rescale the recycling inflow is active (nstep > freq) and KE is computed (cnt > freq)
    assert d["ke"] > 0 and d["guard"] == 0
    assert 0.2 < s.lamda < 5.0 and s.bltkns.max() > 0
