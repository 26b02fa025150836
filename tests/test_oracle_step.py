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


# ----------------------------------------------------------------------------
# The composition of the blocks (dnssetbl.f90:253-1262), pinned by what the
# reference's block order implies for the step's inputs and outputs.
#
# Order read from the reference: time/nstep/cnt += (:254-256); Utold = Ut
# (:257); recycling inflow when nstep > freqTimAvg (:264); convection +
# old-pressure gradient -> Q (:306-369); Gauss-Seidel diffusion (:371-433);
# pRhs (:437-447); GMRES (:449-671); PCorr faces, Press += PCorr, projection
# (:673-726); velocity then pressure BCs (:728-797); divergence (:851-882);
# mass flux (:885-904); statistics (:936-984); rescale when
# mod(nstep, freqTimAvg) == 0 (:986-1220); KE + guard when cnt > freqTimAvg
# (:1223-1261).
# ----------------------------------------------------------------------------
TIGHT_STEP = dict(errorstop=1e-13, gs_mmax=4000, errtol=1e-12, kmax=30, mmax=2000)


def _snapshot(s):
    keys = ("Ut", "Press", "PCorr", "sums", "avg", "Wt", "newk", "Ume", "Ufluc_rec", "bltkns")
    snap = {k: np.array(getattr(s, k), copy=True, order="F") for k in keys}
    snap.update(lamda=s.lamda, lamda03=s.lamda03, nstep=s.nstep, cnt=s.cnt)
    return snap


def _grad(g, PC):
    """central gradient of PCorr at the interior nodes as :711-726 forms it"""
    nX, nY, nZ = g.nX, g.nY, g.nZ
    jm1 = np.r_[nY - 2, np.arange(1, nY - 2)]                      # 0-based tables over j = 1..nY-2
    jp1 = np.r_[np.arange(2, nY - 1), 1]
    Ji = g.mesh["JacInv"][1:-1]
    g1 = 0.5 * (PC[2:, 1:-1, 1:-1] - PC[:-2, 1:-1, 1:-1]) * (1.0 / g.dx)
    g2 = 0.5 * (PC[1:-1, jp1, 1:-1] - PC[1:-1, jm1, 1:-1]) * (1.0 / g.dy)
    g3 = 0.5 * (PC[1:-1, 1:-1, 2:] - PC[1:-1, 1:-1, :-2]) * ((1.0 / g.dXi) * Ji)[None, None, :]
    return g1, g2, g3


_ON = {}


def _always_on(g):
    """the same grid with freqTimAvg = 1: its recycling block is active for
    every nstep >= 2"""
    key = id(g)
    if key not in _ON:
        _ON[key] = (g, orc.Grid(g.nX, g.nY, g.nZ, P=g.P, i_rec=g.i_rec, planes=g.planes, freq_tim_avg=1))
    return _ON[key][1]


def _invariants(g, s0, s, d, helm_tol=1e-10):
    """Failures of the relations one step from snapshot s0 to state s (after
    the step, diagnostics d) must satisfy under the reference's block order;
    an empty list when all hold."""
    bad = []
    I = (slice(1, -1),) * 3
    n1, c1 = s0["nstep"] + 1, s0["cnt"] + 1
    # counters and Utold = the velocity before the step (:254-257)
    if (s.nstep, s.cnt) != (n1, c1):
        bad.append("counters")
    if not np.array_equal(s.Utold[0], s0["Ut"][0]):
        bad.append("Utold is not the pre-step velocity at the inflow plane")
    # recycling inflow exactly when nstep > freqTimAvg (:264), with the
    # incremented nstep in eps (:259-262); later blocks leave the interior of
    # plane i = 1 alone
    # (the recycled plane comes from a twin grid whose block is always on, so
    # the switch step itself is checked here, not inside the block)
    Urec = s0["Ut"].copy(order="F")
    if n1 > g.freq_tim_avg:
        _always_on(g).recycle_inflow(Urec, n1, s0["lamda"], s0["avg"], s0["Wt"], s0["newk"], s0["Ume"],
                                     s0["Ufluc_rec"])
    if not np.array_equal(s.Ut[0, 1:-1, 1:-1], Urec[0, 1:-1, 1:-1]):
        bad.append("inflow plane")
    # Q from the recycled velocity and the old pressure (:306-369)
    _, Qw = g.momentum_rhs(Urec, s0["Press"])
    if not np.array_equal(s.Q[I], Qw[I]):
        bad.append("Q is not momentum_rhs(recycled velocity, old pressure)")
    # undo the projection: u* = U + dt G(PCorr) in the interior, the
    # pre-projection boundary values elsewhere (the BCs come later)
    g1, g2, g3 = _grad(g, s.PCorr)
    us = Urec.copy(order="F")
    for p, gp in enumerate((g1, g2, g3)):
        us[1:-1, 1:-1, 1:-1, p] = s.Ut[1:-1, 1:-1, 1:-1, p] + g.dt * gp
    # the diffusion solve: u* solves the Helmholtz systems with this Q (:371-433)
    for p in range(3):
        r = g.helm_residual(s.Q, us, p + 1)          # 1-based component, as the reference's p
        if not r <= helm_tol * max(np.abs(s.Q[..., p][I]).sum(), 1e-300):
            bad.append("u* does not solve the Helmholtz system (component %d)" % p)
    # pRhs of u* (:437-447), then GMRES to its tolerance (:449-671)
    pr = g.ppe_rhs(us)
    if not np.abs(pr[I] - s.pRhs[I]).max() <= 1e-9 * max(np.abs(s.pRhs[I]).max(), 1e-300):
        bad.append("pRhs is not the divergence of u*")
    # (the solve runs with zero PCorr faces; the projection extrapolates them afterwards)
    pc = np.zeros_like(s.PCorr)
    pc[I] = s.PCorr[I]
    res = g.poisson_apply(pc)[I] - s.pRhs[I]
    if not np.abs(res).max() <= 1e-6 * np.abs(s.pRhs[I]).max():
        bad.append("PCorr does not solve A PCorr = pRhs")
    # Press += PCorr on the interior with the old pressure (:711-714)
    if not np.array_equal(s.Press[I], s0["Press"][I] + s.PCorr[I]):
        bad.append("Press is not old Press + PCorr")
    # BCs last, from the projected velocity and the pre-step Utold: applying
    # them again changes nothing (:728-797)
    U2, P2 = s.Ut.copy(order="F"), s.Press.copy(order="F")
    g.apply_bcs(U2, s0["Ut"], P2, s.wall_w)
    if not (np.array_equal(U2, s.Ut) and np.array_equal(P2, s.Press)):
        bad.append("boundary conditions are not applied to the final fields with the pre-step Utold")
    # diagnostics of the final field (:851-904)
    if d["div"] != g.div_check(s.Ut)[0] or not np.array_equal(d["mass"], g.mass_flux(s.Ut)):
        bad.append("divergence / mass flux not of the final field")
    # statistics of the final field, rescale after them (:936-1220)
    sums = s0["sums"].copy(order="F")
    g.stats_accumulate(s.Ut, sums)
    if n1 % g.freq_tim_avg == 0:
        r = g.rescale(sums, g.freq_tim_avg, s.Ut, s0["bltkns"])
        if not (d["rescaled"] and np.array_equal(s.avg, r["avg"]) and s.lamda == r["lamda"]
                and np.array_equal(s.newk, r["newk"]) and np.array_equal(s.Wt, r["Wt"])
                and not np.any(s.sums)):
            bad.append("rescale is not of this step's accumulated statistics")
        if not np.array_equal(s.Ufluc_rec[:, 1:-1], r["Ufluc_rec"][:, 1:-1]):
            bad.append("recycling fluctuations not refreshed by the rescale")
    else:
        if d["rescaled"] or not np.array_equal(s.sums, sums):
            bad.append("statistics are not the final field's")
    # KE of the final field against the current averages, when cnt > freq (:1223)
    if c1 > g.freq_tim_avg:
        if d["ke"] != g.ke_guard(s.Ut, s.avg)[0]:
            bad.append("KE is not of the final field and the current averages")
    elif d["ke"] != 0.0:
        bad.append("KE computed before cnt > freqTimAvg")
    return bad


def _switch_case():
    """freqTimAvg = 3 from nstep = cnt = 1: steps 2..6 cross a rescale
    without recycling (nstep 3), the first recycled inflow and KE (nstep 4),
    and a rescale in a step with KE on (nstep 6)."""
    g, s = make_state(nX=16, nY=8, nZ=12, freq=3, seed=5, nstep0=1)
    s.cnt = 1
    # a recycling plane whose statistics differ from the inflow so that the
    # recycled plane is distinguishable
    s.avg[..., 1] = 0.05 * g.Ue
    s.Ufluc_rec[...] *= 1.7
    return g, s


def test_step_composition_invariants_hold_over_the_switch_steps():
    g, s = _switch_case()
    seen = []
    for _ in range(5):
        s0 = _snapshot(s)
        d = s.step(**TIGHT_STEP)
        assert _invariants(g, s0, s, d) == [], s.nstep
        seen.append((s.nstep, d["rescaled"], d["ke"] > 0))
    assert seen == [(2, False, False), (3, True, False), (4, False, True), (5, False, True), (6, True, True)]


def _composed(s, mutation=None, errorstop=1e-6, gs_mmax=1000, pc=orc.PC_SSOR_BLOCK, errtol=1e-5, kmax=50,
              mmax=1000, gs_order=orc.GS_LEX_BLOCK):
    """The step composed from the oracle's blocks in the reference's order, or
    with one planted misordering (the mutation check)"""
    g = s.g
    m = mutation or ""
    s.time += g.dt
    s.nstep += 1
    s.cnt += 1
    if m != "utold_after_recycle":
        s.Utold[...] = s.Ut
    on = (s.nstep >= g.freq_tim_avg) if m == "recycle_at_freq" else (s.nstep > g.freq_tim_avg)

    def recycle():
        if on:
            _always_on(g).recycle_inflow(s.Ut, s.nstep, s.lamda, s.avg, s.Wt, s.newk, s.Ume, s.Ufluc_rec)

    if m != "momentum_before_recycle":
        recycle()
    if m == "utold_after_recycle":
        s.Utold[...] = s.Ut
    s.conv, s.Q = g.momentum_rhs(s.Ut, s.Press)
    if m == "momentum_before_recycle":
        recycle()
    if m == "ppe_rhs_before_diffusion":
        s.pRhs = g.ppe_rhs(s.Ut)
    g.diffusion_solve(s.Q, s.Ut, gs_order, errorstop, gs_mmax)
    if m != "ppe_rhs_before_diffusion":
        s.pRhs = g.ppe_rhs(s.Ut)
    g.ppe_gmres(s.pRhs, s.PCorr, pc, errtol, kmax, mmax)
    if m == "bcs_before_project":
        g.apply_bcs(s.Ut, s.Utold, s.Press, s.wall_w)
    g.project(s.PCorr, s.Press, s.Ut)
    if m != "bcs_before_project":
        g.apply_bcs(s.Ut, s.Utold, s.Press, s.wall_w)
    if m == "stats_before_bcs":
        pass
    div, _ = g.div_check(s.Ut)
    mass = g.mass_flux(s.Ut)
    rescaled, ke, guard = False, 0.0, 0
    ke_on = (s.cnt >= g.freq_tim_avg) if m == "ke_at_freq" else (s.cnt > g.freq_tim_avg)
    if m == "ke_before_rescale" and ke_on:
        ke, guard = g.ke_guard(s.Ut, s.avg, s.Ufluc_rec)

    def rescale():
        r = g.rescale(s.sums, g.freq_tim_avg, s.Ut, s.bltkns)
        s.avg, s.lamda, s.lamda03 = r["avg"], r["lamda"], r["lamda03"]
        s.Wt, s.newk, s.bltkns = r["Wt"], r["newk"], r["bltkns"]
        s.deltastar, s.theta = r["deltastar"], r["theta"]
        s.Ufluc_rec[:, 1:-1, :] = r["Ufluc_rec"][:, 1:-1, :]
        s.sums[...] = 0.0

    if m == "rescale_before_stats" and s.nstep % g.freq_tim_avg == 0:
        rescale()
        rescaled = True
    g.stats_accumulate(s.Ut, s.sums)
    if m != "rescale_before_stats" and s.nstep % g.freq_tim_avg == 0:
        rescale()
        rescaled = True
    if m != "ke_before_rescale" and ke_on:
        ke, guard = g.ke_guard(s.Ut, s.avg, s.Ufluc_rec)
    return dict(div=div, mass=mass, ke=ke, guard=guard, rescaled=rescaled)


MUTATIONS = ["utold_after_recycle", "recycle_at_freq", "momentum_before_recycle", "ppe_rhs_before_diffusion",
             "bcs_before_project", "rescale_before_stats", "ke_before_rescale", "ke_at_freq"]


def test_composition_helper_is_the_oracle_step():
    g, s1 = _switch_case()
    _, s2 = _switch_case()
    for _ in range(5):
        d1 = s1.step(**TIGHT_STEP)
        d2 = _composed(s2, None, **TIGHT_STEP)
        assert d1["div"] == d2["div"] and d1["ke"] == d2["ke"] and d1["rescaled"] == d2["rescaled"]
    for k in ("Ut", "Press", "PCorr", "sums", "avg", "Wt", "newk", "Ufluc_rec"):
        assert np.array_equal(getattr(s1, k), getattr(s2, k)), k


@pytest.mark.parametrize("mutation", MUTATIONS)
def test_a_reordered_step_fails_the_invariants(mutation):
    """mutation check: each planted misordering breaks at least one invariant
    on the switch-step run (so the invariants pin the order)"""
    g, s = _switch_case()
    failed = []
    for _ in range(5):
        s0 = _snapshot(s)
        d = _composed(s, mutation, **TIGHT_STEP)
        failed += _invariants(g, s0, s, d)
    assert failed, mutation


def _couette_profile(g):
    """discrete Couette flow of the wall-normal operator: Ab(k) U(k-1) +
    At(k) U(k+1) - (Ab(k) + At(k)) U(k) = 0 on the interior, U = 0 at the
    wall, scaled to reach Ue below the top"""
    Ab, At = g.mesh["Ab"], g.mesh["At"]
    U = np.zeros(g.nZ)
    U[1] = 1.0
    for k in range(1, g.nZ - 1):
        U[k + 1] = ((Ab[k] + At[k]) * U[k] - Ab[k] * U[k - 1]) / At[k]
    return U * (g.Ue / U[g.nZ - 2])


def test_steady_shear_step_closed_form():
    """u = U(z) the discrete Couette profile, v = w = P = PCorr = 0, no
    recycling (nstep <= freqTimAvg): convection and the pressure gradient
    vanish, U solves its own implicit-Euler step, pRhs = 0 and PCorr = 0;
    the step changes only the top value, to the zero-gradient extrapolation
    (4U(nZ-1) - U(nZ-2))/3 (:731), and the boundary pressures, which reduce
    to: bottom 0 (w = 0, :738-755), top 0 (:757-765), outflow 0 (:767-771),
    inflow 2 dx (-rt6/Re)/3 with rt6 the stretched-grid d2U/dz2 of
    :773-794 and (u - Utold)/dt = 0."""
    g, s = make_state(nX=16, nY=8, nZ=12, freq=500, nstep0=3)
    U = _couette_profile(g)
    s.Ut[...] = 0.0
    s.Ut[..., 0] = U[None, None, :]
    s.Press[...] = 0.0
    s.PCorr[...] = 0.0
    Ut0 = s.Ut.copy(order="F")
    d = s.step(**TIGHT_STEP)
    scale = g.Ue
    want = Ut0.copy(order="F")
    want[:, :, -1, 0] = (4.0 * U[-2] - U[-3]) / 3.0
    assert np.abs(s.Ut - want).max() <= 1e-12 * scale
    assert np.abs(s.PCorr).max() <= 1e-12 * scale
    P = s.Press
    assert np.abs(P[1:-1, 1:-1, 0]).max() <= 1e-10 and np.abs(P[1:-1, 1:-1, -1]).max() <= 1e-10
    assert np.abs(P[-1, 1:-1, 1:-1]).max() <= 1e-10
    dXi, Ji, J2 = g.dXi, g.mesh["JacInv"], g.mesh["Jac2"]
    for k in range(1, g.nZ - 1):                     # 0-based interior k
        if k <= 1:
            d1 = (U[k + 1] - U[k]) / dXi
            d2 = (U[k + 2] - 2.0 * U[k + 1] + U[k]) / dXi ** 2
        elif k >= g.nZ - 2:
            d1 = (U[k] - U[k - 1]) / dXi
            d2 = (U[k - 2] - 2.0 * U[k - 1] + U[k]) / dXi ** 2
        else:
            d1 = (U[k + 1] - U[k - 1]) / (2.0 * dXi)
            d2 = (U[k + 1] - 2.0 * U[k] + U[k - 1]) / dXi ** 2
        rt6 = d2 * Ji[k] ** 2 - J2[k] * d1 * Ji[k] ** 3
        want_p = 2.0 * g.dx * (-rt6 / g.Re) / 3.0
        assert P[0, 1:-1, k] == pytest.approx(want_p, rel=1e-9, abs=1e-10), k
    assert d["ke"] == 0.0 and not d["rescaled"]
