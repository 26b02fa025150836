"""GPU parity: every block of the step through the C ABI against the oracle.

FP64 stencil blocks keep the reference's arithmetic order and are compared
bit for bit; reductions (divergence sum, KE, residual norms) differ only in
summation order and are compared at 1e-12 relative to the sum of magnitudes;
the iterative solves are compared (a) against the oracle running the same
red-black orderings, where iteration counts must agree, and (b) against the
reference's lexicographic per-slab orderings at tight tolerance, where only
the converged solution is unique.
"""
import numpy as np
import pytest

from oracle import orc
from tests.case_util import C1_GRID, RAGGED, REF_GRID, gpu_from_case, make_case, ulps

pytestmark = pytest.mark.gpu

GRIDS = [RAGGED, C1_GRID, REF_GRID]


def _fields(sv):
    Ut, P, PC = sv.download()
    return Ut, P, PC


@pytest.fixture(scope="module")
def tbl():
    from turbulentbl_amd import tbl as t
    t.load()
    return t


@pytest.mark.parametrize("grid", GRIDS)
def test_mesh_is_the_oracles(tbl, grid):
    g, s = make_case(grid)
    with gpu_from_case(g, s) as sv:
        m = sv.mesh()
    for k in ("x1", "x2", "x3", "Jac", "Jac2", "Ab", "At", "Ap"):
        assert np.array_equal(m[k], g.mesh[k]), k


@pytest.mark.parametrize("grid", GRIDS)
def test_momentum_rhs_bitwise(tbl, grid):
    g, s = make_case(grid, seed=3)
    rng = np.random.default_rng(0)
    s.Press[...] = rng.standard_normal(s.Press.shape)
    with gpu_from_case(g, s) as sv:
        sv.upload(Press=s.Press)
        sv.momentum_rhs()
        Q = sv.get_field(tbl.F_Q)
    _, Qo = g.momentum_rhs(s.Ut, s.Press)
    I = (slice(1, -1),) * 3
    assert ulps(Q[I], Qo[I]) == 0
    # both Peclet branches were exercised (x, y or z)
    U = s.Ut[2:-2, 2:-2, 2:-2]
    dz = np.diff(g.mesh["x3"])[2:g.nZ - 2]
    Pe = np.concatenate([np.abs(g.Re * U[..., 0] * g.dx).ravel(), np.abs(g.Re * U[..., 1] * g.dy).ravel(),
                         np.abs(g.Re * U[..., 2] * dz[None, None, :]).ravel()])
    assert (Pe > 2).any() and (Pe <= 2).any()


@pytest.mark.parametrize("grid", GRIDS + [(140, 12, 40), (70, 8, 20), (130, 38, 24)])
def test_momentum_rhs_mixed_within_fp32_rounding(tbl, grid):
    """the FP32 momentum kernel (fused multiply-adds; plane tiles staged in LDS
    with periodic span halos, partial edge tiles in x and y) against the
    oracle's FP64 Q from the same FP32-rounded inputs: within 1e-5 of max|Q|
    at every point whose Peclet numbers are not within 1e-4 of the switch at 2
    (there FP32 and FP64 may pick different forms)"""
    g, s = make_case(grid, seed=3)
    rng = np.random.default_rng(0)
    s.Press[...] = rng.standard_normal(s.Press.shape)
    Ut32 = np.asfortranarray(s.Ut.astype(np.float32).astype(np.float64))
    P32 = np.asfortranarray(s.Press.astype(np.float32).astype(np.float64))
    with gpu_from_case(g, s, precision="mixed") as sv:
        sv.upload(Press=s.Press)
        sv.momentum_rhs()
        Q = sv.get_field(tbl.F_Q)
    _, Qo = g.momentum_rhs(Ut32, P32)
    I = (slice(1, -1),) * 3
    d = np.abs(Q[I] - Qo[I])
    U = Ut32[1:-1, 1:-1, 1:-1]
    dz = np.diff(g.mesh["x3"])[1:g.nZ - 1]
    near = np.zeros(d.shape[:3], bool)
    for comp, h in ((0, g.dx), (1, g.dy)):
        near |= np.abs(np.abs(g.Re * U[..., comp] * h) - 2) < 1e-4
    near |= np.abs(np.abs(g.Re * U[..., 2] * dz[None, None, :]) - 2) < 1e-4
    scale = np.abs(Qo[I]).max()
    assert scale > 0
    assert d[~near].max() <= 1e-5 * scale


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("op,pc", [(0, None), (1, orc.PC_SSOR_RB)])
def test_poisson_and_rb_ssor_bitwise(tbl, grid, op, pc):
    g, s = make_case(grid)
    rng = np.random.default_rng(5)
    x = g.zeros()
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    with gpu_from_case(g, s, precond=tbl.TBL_PC_SSOR_RB) as sv:
        y = sv.apply_operator(op, x)
    yo = g.poisson_apply(x) if op == 0 else g.precond_apply(x, pc)
    I = (slice(1, -1),) * 3
    assert ulps(y[I], yo[I]) == 0


def test_jacobi_preconditioner(tbl):
    g, s = make_case(RAGGED)
    x = g.zeros()
    x[1:-1, 1:-1, 1:-1] = np.random.default_rng(6).standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    with gpu_from_case(g, s, precond=tbl.TBL_PC_JACOBI) as sv:
        y = sv.apply_operator(1, x)
    assert ulps(y[1:-1, 1:-1, 1:-1], g.precond_apply(x, orc.PC_JACOBI)[1:-1, 1:-1, 1:-1]) == 0


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("mode,order", [(0, orc.GS_RB), (1, orc.GS_ZLINE_RB), (3, orc.GS_ZEBRA)])
def test_diffusion_gs_matches_oracle_same_ordering(tbl, grid, mode, order):
    g, s = make_case(grid, seed=2)
    _, Q = g.momentum_rhs(s.Ut, s.Press)
    with gpu_from_case(g, s, helm_mode=mode) as sv:
        sv.set_field(tbl.F_Q, Q)
        sw, res = sv.diffusion_solve(1e-6, 1000)
        Ut = sv.get_field(tbl.F_UT)
    Uo = s.Ut.copy(order="F")
    swo, reso = g.diffusion_solve(Q, Uo, order=order, errorstop=1e-6, mmax=1000)
    assert list(sw) == list(swo)
    assert ulps(Ut, Uo) == 0
    assert np.allclose(res, reso, rtol=1e-12)


@pytest.mark.parametrize("mode,order", [(1, orc.GS_ZLINE_RB), (3, orc.GS_ZEBRA)])
@pytest.mark.parametrize("grid", [(20, 10, 192), (12, 8, 300), (140, 12, 40)])
def test_zline_gs_deep_columns_bitwise(tbl, grid, mode, order):
    """z-line sweeps on columns of 190 (12 chunks of <= 16 planes) and 298
    (10 chunks of <= 32 planes) interior planes, as in the 1024x192x256 and
    2048x256x512 cases; and rows of 138 interior columns (three 64-column
    tiles per row, the last one partial, for the zebra ordering)"""
    g, s = make_case(grid, seed=7)
    _, Q = g.momentum_rhs(s.Ut, s.Press)
    with gpu_from_case(g, s, helm_mode=mode) as sv:
        sv.set_field(tbl.F_Q, Q)
        sw, res = sv.diffusion_solve(1e-10, 50)
        Ut = sv.get_field(tbl.F_UT)
    Uo = s.Ut.copy(order="F")
    swo, reso = g.diffusion_solve(Q, Uo, order=order, errorstop=1e-10, mmax=50)
    assert list(sw) == list(swo)
    assert ulps(Ut, Uo) == 0


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("mode", [0, 1, 3])
def test_diffusion_converges_to_reference_lexicographic_solution(tbl, grid, mode):
    g, s = make_case(grid, seed=2)
    _, Q = g.momentum_rhs(s.Ut, s.Press)
    with gpu_from_case(g, s, helm_mode=mode) as sv:
        sv.set_field(tbl.F_Q, Q)
        sv.diffusion_solve(1e-14, 2000)
        Ut = sv.get_field(tbl.F_UT)
    Uo = s.Ut.copy(order="F")
    g.diffusion_solve(Q, Uo, order=orc.GS_LEX_BLOCK, errorstop=1e-14, mmax=2000)
    assert np.abs(Ut - Uo).max() <= 1e-11 * np.abs(Uo).max()


@pytest.mark.parametrize("grid", GRIDS)
def test_ppe_rhs_bitwise(tbl, grid):
    g, s = make_case(grid, seed=4)
    with gpu_from_case(g, s) as sv:
        sv.ppe_rhs()
        r = sv.get_field(tbl.F_PRHS)
    assert ulps(r[1:-1, 1:-1, 1:-1], g.ppe_rhs(s.Ut)[1:-1, 1:-1, 1:-1]) == 0


def _prhs_case(g, s):
    pRhs = g.ppe_rhs(s.Ut)
    PC = g.zeros()
    PC[1:-1, 1:-1, 1:-1] = np.random.default_rng(9).standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2)) * 1e-3
    return pRhs, PC


@pytest.mark.parametrize("grid", GRIDS)
def test_gmres_matches_oracle_same_preconditioner(tbl, grid):
    g, s = make_case(grid, seed=4)
    pRhs, PC0 = _prhs_case(g, s)
    kmax, errtol = 30, 1e-8
    with gpu_from_case(g, s, kmax=kmax, precond=tbl.TBL_PC_SSOR_RB) as sv:
        sv.set_field(tbl.F_PRHS, pRhs)
        sv.set_field(tbl.F_PCORR, PC0)
        st = sv.ppe_solve(errtol, kmax, 1000)
        PC = sv.get_field(tbl.F_PCORR)
    PCo = PC0.copy(order="F")
    sto = g.ppe_gmres(pRhs, PCo, pc=orc.PC_SSOR_RB, errtol=errtol, kmax=kmax, mmax=1000)
    assert st["cycles"] == sto["cycles"]
    assert abs(st["iters"] - sto["iters"]) <= 1
    # agreement at the level the stopping tolerance leaves open
    diff = np.abs(PC - PCo).max() / np.abs(PCo).max()
    assert diff <= 1e-6
    assert st["true_res"] == pytest.approx(sto["true_res"], rel=0.05)


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("pc", [1, 4])
def test_gmres_tight_matches_reference_block_ssor_solution(tbl, grid, pc):
    g, s = make_case(grid, seed=4)
    pRhs, PC0 = _prhs_case(g, s)
    with gpu_from_case(g, s, kmax=40, precond=pc) as sv:
        sv.set_field(tbl.F_PRHS, pRhs)
        sv.set_field(tbl.F_PCORR, PC0)
        st = sv.ppe_solve(1e-12, 40, 2000)
        PC = sv.get_field(tbl.F_PCORR)
    if pc == tbl.TBL_PC_FAST:      # the separable preconditioner is A^{-1} up to rounding
        assert st["iters"] <= 6
    PCo = PC0.copy(order="F")
    g.ppe_gmres(pRhs, PCo, pc=orc.PC_SSOR_BLOCK, errtol=1e-12, kmax=40, mmax=2000)
    assert np.abs(PC - PCo).max() <= 1e-9 * np.abs(PCo).max()
    # faces zeroed by the solver
    assert np.all(PC[0] == 0) and np.all(PC[:, :, -1] == 0)


@pytest.mark.parametrize("grid", GRIDS)
def test_project_and_bcs_bitwise(tbl, grid):
    g, s = make_case(grid, seed=5)
    rng = np.random.default_rng(2)
    PC = g.zeros()
    PC[1:-1, 1:-1, 1:-1] = rng.standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2)) * 1e-2
    Press = np.asfortranarray(rng.standard_normal(s.Press.shape))
    Utold = np.asfortranarray(s.Ut + 1e-3 * rng.standard_normal(s.Ut.shape))
    with gpu_from_case(g, s) as sv:
        sv.upload(Press=Press, PCorr=PC)
        sv.set_field(tbl.F_UTOLD, Utold)
        sv.project()
        sv.apply_bcs()
        Ut, P, PCg = sv.download()
    Uo, Po, PCo = s.Ut.copy(order="F"), Press.copy(order="F"), PC.copy(order="F")
    g.project(PCo, Po, Uo)
    g.apply_bcs(Uo, Utold, Po)
    assert ulps(PCg, PCo) == 0
    assert ulps(Ut, Uo) == 0
    assert ulps(P, Po) == 0


@pytest.mark.parametrize("grid", GRIDS)
def test_diagnostics(tbl, grid):
    g, s = make_case(grid, seed=6)
    with gpu_from_case(g, s) as sv:
        div, sdiv = sv.div_check()
        mass = sv.mass_flux()
        sv.stats_accumulate()
        sv.stats_accumulate()
        st = sv.get_state()
        ke, gd = sv.ke_guard()
        uf = sv.get_state()["Ufluc_rec"]
        tau = sv.wall_shear()
    divo, sdivo = g.div_check(s.Ut)
    # signed sum: tolerance relative to the sum of magnitudes of its terms
    mag = np.abs(g.ppe_rhs(s.Ut)).sum() * g.dt * 10 + 1.0
    assert abs(sdiv - sdivo) <= 1e-12 * mag
    assert np.allclose(mass, g.mass_flux(s.Ut), rtol=1e-13)
    sums = g.zeros2(9)
    g.stats_accumulate(s.Ut, sums)
    g.stats_accumulate(s.Ut, sums)
    assert ulps(st["sums"], sums) == 0
    ufo = s.Ufluc_rec.copy(order="F")
    keo, gdo = g.ke_guard(s.Ut, s.avg, ufo)
    assert ke == pytest.approx(keo, rel=1e-12) and gd == gdo == 0
    assert ulps(uf[1:-1, 1:-1], ufo[1:-1, 1:-1]) == 0
    assert ulps(tau, g.wall_shear(s.Ut)) == 0


@pytest.mark.parametrize("grid", [C1_GRID, REF_GRID])
def test_rescale(tbl, grid):
    g, s = make_case(grid, seed=7, freq=3)
    sums = g.zeros2(9)
    for q in range(3):       # three slightly different fields, like consecutive steps
        g.stats_accumulate(np.asfortranarray(s.Ut * (1 + 1e-3 * q)), sums)
    with gpu_from_case(g, s) as sv:
        sv.set_state(sums=sums, bltkns=s.bltkns, avg=s.avg, Wt=s.Wt, newk=s.newk, Ume=s.Ume,
                     Ufluc_rec=s.Ufluc_rec, lamda=1.0, lamda03=1.0, nstep=s.nstep)
        sv.rescale()
        st = sv.get_state()
    r = g.rescale(sums, 3, s.Ut, s.bltkns)
    assert ulps(st["avg"], r["avg"]) == 0
    assert not st["sums"].any()
    assert st["lamda"] == r["lamda"] and st["lamda03"] == r["lamda03"]
    assert np.array_equal(st["bltkns"], r["bltkns"])
    assert np.array_equal(st["newk"], r["newk"])
    assert np.allclose(st["Wt"], r["Wt"], rtol=1e-14, atol=1e-15)   # device tanh vs libm tanh
    assert ulps(st["deltastar"], r["deltastar"]) == 0 and ulps(st["theta"], r["theta"]) == 0
    assert ulps(st["Ufluc_rec"][1:-1, 1:-1], r["Ufluc_rec"][1:-1, 1:-1]) == 0


@pytest.mark.parametrize("grid", GRIDS)
def test_recycle_inflow_bitwise(tbl, grid):
    g, s = make_case(grid, seed=8, nstep0=1234)
    s.newk[3] = min(g.nZ - 1, 7)
    with gpu_from_case(g, s) as sv:
        sv.set_state(avg=s.avg, Wt=s.Wt, newk=s.newk, Ume=s.Ume, Ufluc_rec=s.Ufluc_rec, lamda=1.1,
                     lamda03=1.0, nstep=s.nstep)
        sv.begin_step()
        sv.recycle_inflow()
        Ut = sv.get_field(tbl.F_UT)
    Uo = s.Ut.copy(order="F")
    g.recycle_inflow(Uo, s.nstep + 1, 1.1, s.avg, s.Wt, s.newk, s.Ume, s.Ufluc_rec)
    assert ulps(Ut, Uo) == 0


# ---------------------------------------------------------------- whole steps
def _oracle_steps(s, n, **kw):
    return [s.step(**kw) for _ in range(n)]


@pytest.mark.parametrize("grid", [RAGGED, C1_GRID])
def test_momentum_rhs_dnstbl_scheme_bitwise(tbl, grid):
    """TBL_SCHEME_DNSTBL (dnstbl/dnstbl.f90:620-766) against the oracle's
    variant, 0 ulp in FP64, with a body force; every branch of the Peclet
    window and the sign(1, ceiling(u)) quirk is exercised"""
    g, s = make_case(grid, seed=7, scheme=1, forc=(0.4, -0.1, 0.02))
    rng = np.random.default_rng(2)
    s.Press[...] = rng.standard_normal(s.Press.shape)
    s.Ut[..., 1] += rng.uniform(-1.5, 1.5, s.Ut[..., 1].shape)      # v spans (-1, 0) and <= -1
    with gpu_from_case(g, s) as sv:
        sv.upload(Ut=s.Ut, Press=s.Press)
        sv.momentum_rhs()
        Q = sv.get_field(tbl.F_Q)
    _, Qo = g.momentum_rhs(s.Ut, s.Press)
    I = (slice(1, -1),) * 3
    assert ulps(Q[I], Qo[I]) == 0
    U = s.Ut[2:-2, 2:-2, 2:-2]
    sg = np.where(np.ceil(U) < 0, -1.0, 1.0)
    Pe = np.concatenate([(sg[..., 0] * g.Re * U[..., 0] * g.dx).ravel(), (sg[..., 1] * g.Re * U[..., 1] * g.dy).ravel()])
    assert (Pe <= 2).any() and ((Pe > 2) & (Pe < 50)).any() and (Pe >= 50).any()
    assert ((U[..., 1] > -1) & (U[..., 1] < 0) & (np.abs(g.Re * U[..., 1] * g.dy) > 2)).any()


def test_step_dnstbl_scheme_matches_oracle(tbl):
    """whole steps with the dnstbl convection and a body force: FP64 with the
    same orderings as the oracle, and mixed precision (fast-preconditioned
    FGMRES to a true residual of 1e-10) within the mixed-vs-FP64 bound"""
    g, s = make_case(C1_GRID, seed=1, scheme=1, forc=(0.5, 0.0, 0.0))
    kw = dict(errorstop=1e-6, gs_mmax=1000, errtol=1e-9, kmax=30, mmax=1000)
    with gpu_from_case(g, s, kmax=30, errtol=1e-9, helm_mode=0, precond=tbl.TBL_PC_SSOR_RB) as sv:
        dg = sv.step(3)
        Ut, P, PC = sv.download()
    with gpu_from_case(g, s, precision="mixed", kmax=40, errtol=1e-10, tol_mode=tbl.TBL_TOL_TRUE,
                       inner_tol=1e-3) as sv:
        sv.step(3)
        Um, _, _ = sv.download()
    do = _oracle_steps(s, 3, gs_order=orc.GS_RB, pc=orc.PC_SSOR_RB, **kw)
    for a, b in zip(dg, do):
        assert a["gs_sweeps"] == list(b["gs_sweeps"])
    assert np.abs(Ut - s.Ut).max() <= 1e-9 * np.abs(s.Ut).max()
    assert np.abs(Um - s.Ut).max() <= 2e-6 * np.abs(s.Ut).max()


@pytest.mark.parametrize("grid", [RAGGED, C1_GRID])
def test_step_matches_oracle_same_orderings(tbl, grid):
    g, s = make_case(grid, seed=1)
    kw = dict(errorstop=1e-6, gs_mmax=1000, errtol=1e-9, kmax=30, mmax=1000)
    with gpu_from_case(g, s, kmax=30, errtol=1e-9, helm_mode=0, precond=tbl.TBL_PC_SSOR_RB) as sv:
        dg = sv.step(3)
        Ut, P, PC = sv.download()
    do = _oracle_steps(s, 3, gs_order=orc.GS_RB, pc=orc.PC_SSOR_RB, **kw)
    for a, b in zip(dg, do):
        assert a["gs_sweeps"] == list(b["gs_sweeps"])
        assert abs(a["gm_iters"] - b["gmres"]["iters"]) <= 2
        assert np.allclose(a["mass"], b["mass"], rtol=1e-10)
    assert np.abs(Ut - s.Ut).max() <= 1e-9 * np.abs(s.Ut).max()
    Pi = (slice(1, -1),) * 3
    assert np.abs(P[Pi] - s.Press[Pi]).max() <= 1e-7 * np.abs(s.Press[Pi]).max()


@pytest.mark.parametrize("grid", [RAGGED, C1_GRID])
def test_step_matches_reference_orderings_at_tight_tolerance(tbl, grid):
    g, s = make_case(grid, seed=1)
    kw = dict(errorstop=1e-13, gs_mmax=3000, errtol=1e-12, kmax=40, mmax=2000)
    with gpu_from_case(g, s, kmax=40, errtol=1e-12, errorstop=1e-13, gs_mmax=3000, mmax=2000) as sv:
        sv.step(2)
        Ut, P, PC = sv.download()
    _oracle_steps(s, 2, gs_order=orc.GS_LEX_BLOCK, pc=orc.PC_SSOR_BLOCK, **kw)
    assert np.abs(Ut - s.Ut).max() <= 1e-9 * np.abs(s.Ut).max()


def test_step_rescale_path_and_determinism(tbl):
    g, s = make_case(RAGGED, freq=3, nstep0=0)
    outs = []
    for _ in range(2):
        with gpu_from_case(g, s, helm_mode=0, precond=tbl.TBL_PC_SSOR_RB) as sv:
            d = sv.step(5)
            outs.append((sv.download(), sv.get_state(), d))
    (f1, st1, d1), (f2, st2, d2) = outs
    for a, b in zip(f1, f2):
        assert np.array_equal(a, b)
    assert [x["rescaled"] for x in d1] == [0, 0, 1, 0, 0]
    assert d1[-1]["ke"] > 0 and d1[-1]["ke"] == d2[-1]["ke"]
    do = _oracle_steps(s, 5, gs_order=orc.GS_RB, pc=orc.PC_SSOR_RB, errtol=1e-5, kmax=50)
    assert st1["lamda"] == pytest.approx(s.lamda, rel=1e-6)
    assert np.array_equal(st1["newk"], s.newk)
    assert d1[-1]["ke"] == pytest.approx(do[-1]["ke"], rel=1e-5)


def test_failed_step_inside_a_multi_step_call(tbl):
    """A step that trips the guard inside a multi-step call stops the call
    there: the fields, Utold and Q are bit for bit those of the same step
    called alone (no part of the next step has run)."""
    g, s = make_case(RAGGED, freq=1, nstep0=0)
    s.cnt = 5
    from turbulentbl_amd.tbl import TblError
    Ut = s.Ut.copy(order="F")
    Ut[5, 4, 5, 1] += 400.0          # survives one step as a large fluctuation
    out = []
    for n in (1, 3):
        with gpu_from_case(g, s) as sv:
            sv.upload(Ut=Ut)
            with pytest.raises(TblError) as e:
                sv.step(n)
            assert e.value.code == -2
            out.append([sv.get_field(f) for f in (tbl.F_UT, tbl.F_UTOLD, tbl.F_Q, tbl.F_PRESS, tbl.F_PCORR)])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_blowup_guard_and_nan(tbl):
    g, s = make_case(RAGGED, freq=1, nstep0=0)
    s.cnt = 5
    from turbulentbl_amd.tbl import TblError
    Ut = s.Ut.copy(order="F")
    Ut[5, 4, 5, 1] += 400.0          # survives one step as a large fluctuation
    with gpu_from_case(g, s) as sv:
        sv.upload(Ut=Ut)
        with pytest.raises(TblError) as e:
            sv.step(1)
        assert e.value.code == -2
    Ut = s.Ut.copy(order="F")
    Ut[5, 4, 5, 0] = np.nan
    with gpu_from_case(g, s) as sv:
        sv.upload(Ut=Ut)
        with pytest.raises(TblError) as e:
            sv.step(1)
        assert e.value.code in (-3, -2)


def test_noconv_reported(tbl):
    g, s = make_case(RAGGED)
    pRhs, PC0 = _prhs_case(g, s)
    from turbulentbl_amd.tbl import TblError
    with gpu_from_case(g, s, kmax=2, fail_on_noconv=1, precond=tbl.TBL_PC_SSOR_RB) as sv:
        sv.set_field(tbl.F_PRHS, pRhs)
        with pytest.raises(TblError) as e:
            sv.ppe_solve(1e-14, 2, 2)
        assert e.value.code == -4


# ---------------------------------------------------------------- mixed precision
@pytest.mark.parametrize("grid", [C1_GRID])
def test_mixed_precision_step_error_vs_fp64_oracle(tbl, grid):
    g, s = make_case(grid, seed=1)
    kw = dict(errorstop=1e-6, gs_mmax=1000, errtol=1e-10, kmax=40, mmax=1000)
    with gpu_from_case(g, s, precision="mixed", kmax=40, errtol=1e-10, tol_mode=tbl.TBL_TOL_TRUE,
                       inner_tol=1e-3) as sv:
        dm = sv.step(3)
        Ut, P, PC = sv.download()
        tau = sv.wall_shear()
    _oracle_steps(s, 3, gs_order=orc.GS_ZEBRA, pc=orc.PC_SSOR_RB, **kw)
    eu = np.abs(Ut - s.Ut).max() / np.abs(s.Ut).max()
    assert eu <= 2e-6
    # the Gauss-Seidel's first residual comes fused from the momentum kernel
    # (a wrong or NaN first residual leaves it at the sweep cap)
    assert all(max(d["gs_sweeps"]) <= 20 for d in dm)
    tau_o = g.wall_shear(s.Ut)
    cf_err = abs(tau[1:-1].mean() - tau_o[1:-1].mean()) / abs(tau_o[1:-1].mean())
    assert cf_err <= 1e-4
    assert all(d["gm_true_res"] > 0 for d in dm)


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-4)])
def test_fast_preconditioner_inverts_the_poisson_operator(tbl, grid, precision, tol):
    g, s = make_case(grid)
    rng = np.random.default_rng(12)
    x = g.zeros()
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    Ax = g.poisson_apply(x)          # oracle operator, pinned bitwise against the GPU apply
    with gpu_from_case(g, s, precision=precision, precond=tbl.TBL_PC_FAST) as sv:
        y = sv.apply_operator(1, Ax)
    I = (slice(1, -1),) * 3
    assert np.abs(y[I] - x[I]).max() <= tol * np.abs(x[I]).max()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-4)])
def test_fast_preconditioner_wide_transforms(tbl, precision, tol):
    """span and wall-normal transforms wider than 8 column tiles (nY-2 = 298,
    nZ-2 = 268): the matrix-core kernel splits them over column blocks, the
    last one partly past N"""
    g, s = make_case((12, 300, 270))
    rng = np.random.default_rng(13)
    x = g.zeros()
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    Ax = g.poisson_apply(x)
    with gpu_from_case(g, s, precision=precision, precond=tbl.TBL_PC_FAST) as sv:
        y = sv.apply_operator(1, Ax)
    I = (slice(1, -1),) * 3
    assert np.abs(y[I] - x[I]).max() <= tol * np.abs(x[I]).max()


@pytest.mark.parametrize("grid", [RAGGED, C1_GRID, REF_GRID])
def test_default_path_step_matches_oracle_zline_at_tight_tolerance(tbl, grid):
    """library defaults (z-line zebra Gauss-Seidel, fast-preconditioned GMRES)
    against the oracle with the same Gauss-Seidel ordering and block SSOR, both
    solved tightly"""
    g, s = make_case(grid, seed=1)
    with gpu_from_case(g, s, kmax=40, errtol=1e-12, errorstop=1e-13, gs_mmax=400) as sv:
        assert sv.p.helm_mode == tbl.TBL_HELM_ZEBRA
        dg = sv.step(2)
        Ut, P, PC = sv.download()
    do = _oracle_steps(s, 2, gs_order=orc.GS_ZEBRA, pc=orc.PC_SSOR_BLOCK, errorstop=1e-13, gs_mmax=400,
                       errtol=1e-12, kmax=40, mmax=2000)
    for a, b in zip(dg, do):
        assert a["gs_sweeps"] == list(b["gs_sweeps"])
        assert a["gm_iters"] <= 8
    assert np.abs(Ut - s.Ut).max() <= 1e-9 * np.abs(s.Ut).max()


def test_extrapolated_initial_guess_meets_the_same_tolerance(tbl):
    """pc_extrap = 1 (initial guess 2 PCorr(n-1) - PCorr(n-2)) converges to the
    same true residual target; the steps agree with the reference warm start
    to the solve tolerance"""
    g, s = make_case(C1_GRID, seed=9)
    out = []
    for ex in (0, 1):
        with gpu_from_case(g, s, tol_mode=tbl.TBL_TOL_TRUE, errtol=1e-12, pc_extrap=ex) as sv:
            d = sv.step(4)
            out.append((sv.download()[0], d))
    (U0, d0), (U1, d1) = out
    for b in d1:
        assert np.isfinite(b["gm_true_res"]) and b["gm_iters"] >= 0
    assert np.abs(U1 - U0).max() <= 1e-10 * np.abs(U0).max()


@pytest.mark.parametrize("precision,pc", [("fp64", "ssor"), ("mixed", "fast")])
def test_true_residual_stop_uses_interior_norms(tbl, precision, pc):
    """TBL_TOL_TRUE stops on ||pRhs - A PCorr|| <= errtol ||pRhs||, both over the
    interior nodes (the norm of pRhs comes out of the first residual pass).
    Boundary entries of pRhs planted 1e4 times the interior scale must not
    loosen the test; the converged residual is checked with the oracle's
    operator, and the reported true residual is that residual's norm.  The
    preconditioners are ones whose convergence the threshold decides: FP64
    red-black SSOR, and the FP32 fast preconditioner (one cycle reaches ~1e-6)."""
    g, s = make_case(C1_GRID, seed=4)
    rng = np.random.default_rng(21)
    I = (slice(1, -1),) * 3
    b = 1e4 * rng.standard_normal((g.nX, g.nY, g.nZ))
    b[I] = rng.standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    b = np.asfortranarray(b)
    errtol = 1e-8
    precond = tbl.TBL_PC_SSOR_RB if pc == "ssor" else tbl.TBL_PC_FAST
    with gpu_from_case(g, s, precision=precision, precond=precond, tol_mode=tbl.TBL_TOL_TRUE,
                       errtol=errtol, kmax=40) as sv:
        sv.set_field(tbl.F_PCORR, g.zeros())
        sv.set_field(tbl.F_PRHS, b)
        out = sv.ppe_solve(errtol=errtol, kmax=40, mmax=2000)
        x = sv.get_field(tbl.F_PCORR)
    r = b[I] - g.poisson_apply(x)[I]
    rn, bn = np.linalg.norm(r), np.linalg.norm(b[I])
    assert rn <= errtol * bn * (1 + 1e-9)
    assert abs(out["true_res"] - rn) <= 1e-6 * rn


def test_fast_preconditioned_cycles_of_several_arnoldi_steps(tbl):
    """Mixed precision with the fast preconditioner: with the default inner
    tolerance every refinement cycle stops after one Arnoldi step; with an
    inner tolerance one FP32 step cannot reach, cycles run several steps, so
    the kernels queued after a cycle's first step (the cycle's end and the
    next cycle's fused residual pass, gated on the cycle having ended) find it
    still active and must do nothing.  Both solves meet the true-residual
    target, checked with the oracle's operator."""
    g, s = make_case(C1_GRID, seed=4)
    rng = np.random.default_rng(22)
    I = (slice(1, -1),) * 3
    b = g.zeros()
    b[I] = rng.standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    errtol = 1e-9
    got = {}
    for itol in (1e-3, 1e-12):
        with gpu_from_case(g, s, precision="mixed", precond=tbl.TBL_PC_FAST, tol_mode=tbl.TBL_TOL_TRUE,
                           errtol=errtol, kmax=6, inner_tol=itol) as sv:
            sv.set_field(tbl.F_PCORR, g.zeros())
            sv.set_field(tbl.F_PRHS, b)
            out = sv.ppe_solve(errtol=errtol, kmax=6, mmax=200)
            got[itol] = (out, sv.get_field(tbl.F_PCORR))
    bn = np.linalg.norm(b[I])
    for itol, (out, x) in got.items():
        rn = np.linalg.norm(b[I] - g.poisson_apply(x)[I])
        assert rn <= errtol * bn * (1 + 1e-9), (itol, out)
        assert abs(out["true_res"] - rn) <= 1e-6 * rn
    one, many = got[1e-3][0], got[1e-12][0]
    # the cycle counter also counts the final cycle start that finds the target met
    assert one["iters"] == one["cycles"] - 1, one
    assert many["iters"] >= many["cycles"], many


# ---------------------------------------------------------------- a17 window
def test_wall_shear_window_matches_oracle_and_its_own_sums(tbl):
    """tbl_step's wall-shear window (dnstbl.f90:1666-1682): the mean and C_f
    equal, bit for bit, the sum of the per-step tbl_wall_shear values divided
    as the reference divides; against the oracle stepping the same orderings
    they agree to the field tolerance; windows end where nstep % 3 == 0."""
    g, s = make_case(RAGGED, seed=2, nstep0=10)
    s.tau_window = 3
    kw = dict(errorstop=1e-6, gs_mmax=1000, errtol=1e-9, kmax=30, mmax=1000)
    got = []
    with gpu_from_case(g, s, kmax=30, errtol=1e-9, helm_mode=0, precond=tbl.TBL_PC_SSOR_RB,
                       wall_shear_window=3) as sv:
        acc = np.zeros(g.nX)
        for _ in range(7):
            d = sv.step(1)[0]
            tau = sv.wall_shear()
            acc[1:-1] = tau[1:-1] + acc[1:-1]
            if d["cf_ready"]:
                tm, cf, nw = sv.wall_shear_mean()
                sw = acc / 3.0
                assert np.array_equal(tm, sw / (g.nY - 2))
                assert np.array_equal(cf, 2.0 * (sw / (g.Ue * g.Ue * (g.nY - 2))))
                got.append((d["nstep"], tm, cf, nw))
                acc[:] = 0.0
    do = _oracle_steps(s, 7, gs_order=orc.GS_RB, pc=orc.PC_SSOR_RB, **kw)
    assert [n for n, *_ in got] == [12, 15] and got[-1][3] == 2
    assert [d["cf_ready"] for d in do] == [False, True, False, False, True, False, False]
    tm, cf = got[-1][1], got[-1][2]
    assert np.abs(tm - s.taumean).max() <= 1e-8 * np.abs(s.taumean).max()
    assert np.abs(cf - s.cf).max() <= 1e-8 * np.abs(s.cf).max()


def test_wall_shear_mean_needs_a_window(tbl):
    from turbulentbl_amd.tbl import TblError
    g, s = make_case(RAGGED)
    with gpu_from_case(g, s) as sv:
        with pytest.raises(TblError) as e:
            sv.wall_shear_mean()
        assert e.value.code == -5


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_true_residual_solve_of_a_zero_rhs_is_zero(tbl, precision):
    """TBL_TOL_TRUE with pRhs = 0 on the interior and a nonzero initial PCorr:
    the solution of A x = 0 is x = 0, reached in the first cycle (the relative
    test alone could only stop on an exact zero residual)."""
    g, s = make_case(RAGGED, seed=3)
    PC0 = g.zeros()
    PC0[1:-1, 1:-1, 1:-1] = np.random.default_rng(8).standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    with gpu_from_case(g, s, precision=precision, tol_mode=tbl.TBL_TOL_TRUE, errtol=1e-10, kmax=20,
                       mmax=50, fail_on_noconv=1) as sv:
        sv.set_field(tbl.F_PCORR, PC0)
        sv.set_field(tbl.F_PRHS, g.zeros())
        out = sv.ppe_solve(errtol=1e-10, kmax=20, mmax=50)
        x = sv.get_field(tbl.F_PCORR)
    assert out["cycles"] == 1 and out["iters"] == 0
    assert not np.any(x[1:-1, 1:-1, 1:-1])
    assert out["true_res"] == 0.0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_zebra_components_stopping_at_different_sweeps(tbl, precision):
    """The zebra sweep reads one velocity buffer and writes the other; a
    component whose stopping test has ended is copied through.  Components
    that stop after different numbers of sweeps (here v is already converged:
    zero right-hand side and field) must each keep their own result, equal
    to the oracle's, and the field ends in the context's velocity buffer."""
    g, s = make_case(C1_GRID, seed=2)
    _, Q = g.momentum_rhs(s.Ut, s.Press)
    Q[..., 1] = 0.0
    U0 = s.Ut.copy(order="F")
    U0[..., 1] = 0.0
    with gpu_from_case(g, s, precision=precision, helm_mode=3) as sv:
        sv.upload(Ut=U0)
        sv.set_field(tbl.F_Q, Q)
        sw, res = sv.diffusion_solve(1e-9, 50)
        Ut = sv.get_field(tbl.F_UT)
    Uo = U0.copy(order="F")
    swo, reso = g.diffusion_solve(Q, Uo, order=orc.GS_ZEBRA, errorstop=1e-9, mmax=50)
    assert sw[1] == 0 and sw[0] > 0
    if precision == "fp64":
        assert list(sw) == list(swo)
        assert ulps(Ut, Uo) == 0
    else:
        assert np.abs(Ut - Uo).max() <= 1e-5 * np.abs(Uo).max()
    assert not np.any(Ut[..., 1])


@pytest.mark.parametrize("grid", [RAGGED, C1_GRID])
def test_step_zebra_matches_oracle_same_orderings(tbl, grid):
    g, s = make_case(grid, seed=1)
    kw = dict(errorstop=1e-6, gs_mmax=1000, errtol=1e-9, kmax=30, mmax=1000)
    with gpu_from_case(g, s, kmax=30, errtol=1e-9, helm_mode=3, precond=tbl.TBL_PC_SSOR_RB) as sv:
        dg = sv.step(3)
        Ut, P, PC = sv.download()
    do = _oracle_steps(s, 3, gs_order=orc.GS_ZEBRA, pc=orc.PC_SSOR_RB, **kw)
    for a, b in zip(dg, do):
        assert a["gs_sweeps"] == list(b["gs_sweeps"])
    assert np.abs(Ut - s.Ut).max() <= 1e-9 * np.abs(s.Ut).max()
