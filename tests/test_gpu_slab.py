"""GPU parity of the reference's own per-slab orderings against oracle(P).

The reference decomposes the wall-normal direction into P slabs, one per MPI
rank (dnssetbl/dnssetbl.f90:76-92).  Its Gauss-Seidel for the implicit
diffusion is lexicographic inside each slab, with the ghost planes from the
last exchange (:395-430), and its GMRES preconditioner is block SSOR, one
lexicographic forward and backward sweep per slab with no coupling across
slabs or span faces (:509-532, 586-610).  `TBL_HELM_LEX_SLAB` and
`TBL_PC_SSOR_BLOCK` run the same orderings on the GPU (one workgroup per slab,
hyperplane order inside it); both must equal the oracle with the same P bit
for bit in FP64, sweep counts included.  The last tests run the 1024x256x192
configuration split into 8 slabs, one rank's slab at its real shape
(1024x256x24 interior planes plus ghosts).
"""
import numpy as np
import pytest

from oracle import orc
from tests.case_util import C1_GRID, RAGGED, REF_GRID, gpu_from_case, make_case, ulps

pytestmark = pytest.mark.gpu

CASES = [(RAGGED, 1), (RAGGED, 2), (RAGGED, 4), (C1_GRID, 1), (C1_GRID, 3), (C1_GRID, 8), (REF_GRID, 2),
         (REF_GRID, 5), ((140, 12, 40), 3)]


@pytest.fixture(scope="module")
def tbl():
    from turbulentbl_amd import tbl as t
    t.load()
    return t


@pytest.mark.parametrize("grid,P", CASES)
def test_slab_lexicographic_gs_bitwise(tbl, grid, P):
    g, s = make_case(grid, seed=2, P=P)
    _, Q = g.momentum_rhs(s.Ut, s.Press)
    with gpu_from_case(g, s, helm_mode=tbl.TBL_HELM_LEX_SLAB, slabs=P) as sv:
        sv.set_field(tbl.F_Q, Q)
        sw, res = sv.diffusion_solve(1e-6, 1000)
        Ut = sv.get_field(tbl.F_UT)
    Uo = s.Ut.copy(order="F")
    swo, reso = g.diffusion_solve(Q, Uo, order=orc.GS_LEX_BLOCK, errorstop=1e-6, mmax=1000)
    assert list(sw) == list(swo)
    assert ulps(Ut, Uo) == 0
    assert np.allclose(res, reso, rtol=1e-12)


@pytest.mark.parametrize("grid,P", CASES)
def test_block_ssor_bitwise(tbl, grid, P):
    g, s = make_case(grid, P=P)
    rng = np.random.default_rng(5)
    x = g.zeros()
    x[1:-1, 1:-1, 1:-1] = rng.standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    with gpu_from_case(g, s, precond=tbl.TBL_PC_SSOR_BLOCK, slabs=P) as sv:
        y = sv.apply_operator(1, x)
    yo = g.precond_apply(x, orc.PC_SSOR_BLOCK)
    I = (slice(1, -1),) * 3
    assert ulps(y[I], yo[I]) == 0


@pytest.mark.parametrize("grid,P", [(RAGGED, 2), (C1_GRID, 4), (REF_GRID, 5)])
def test_gmres_block_ssor_matches_oracle(tbl, grid, P):
    """the reference's GMRES (preconditioned-residual stopping rule) with its
    block SSOR: the same cycle count, at most one Arnoldi step apart (dot
    products are summed in another order), the solution within what the
    stopping tolerance leaves open"""
    g, s = make_case(grid, seed=4, P=P)
    pRhs = g.ppe_rhs(s.Ut)
    PC0 = g.zeros()
    PC0[1:-1, 1:-1, 1:-1] = np.random.default_rng(9).standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2)) * 1e-3
    kmax, errtol = 30, 1e-8
    with gpu_from_case(g, s, kmax=kmax, precond=tbl.TBL_PC_SSOR_BLOCK, slabs=P) as sv:
        sv.set_field(tbl.F_PRHS, pRhs)
        sv.set_field(tbl.F_PCORR, PC0)
        st = sv.ppe_solve(errtol, kmax, 1000)
        PC = sv.get_field(tbl.F_PCORR)
    PCo = PC0.copy(order="F")
    sto = g.ppe_gmres(pRhs, PCo, pc=orc.PC_SSOR_BLOCK, errtol=errtol, kmax=kmax, mmax=1000)
    assert st["cycles"] == sto["cycles"]
    assert abs(st["iters"] - sto["iters"]) <= 1
    assert np.abs(PC - PCo).max() / np.abs(PCo).max() <= 1e-6


@pytest.mark.parametrize("grid,P", [(C1_GRID, 4), (REF_GRID, 5)])
def test_reference_orderings_whole_step(tbl, grid, P):
    """a whole step with the reference's solvers and orderings (per-slab
    Gauss-Seidel, GMRES(50) with block SSOR and its errtol rule) against the
    oracle with the same P: Gauss-Seidel sweep counts equal, fields within the
    solve tolerances"""
    g, s = make_case(grid, seed=1, P=P)
    with gpu_from_case(g, s, helm_mode=tbl.TBL_HELM_LEX_SLAB, precond=tbl.TBL_PC_SSOR_BLOCK, slabs=P,
                       tol_mode=tbl.TBL_TOL_REF, kmax=50, errtol=1e-5) as sv:
        dg = sv.step(2)
        Ut, Press, PCo_gpu = sv.download()
    do = [s.step() for _ in range(2)]
    for a, b in zip(dg, do):
        assert a["gs_sweeps"] == [int(x) for x in b["gs_sweeps"]]
    assert np.abs(Ut - s.Ut).max() <= 1e-6 * np.abs(s.Ut).max()


# ---------------------------------------------------------------- real slab shape
C3 = (1024, 256, 192)


def _c3_slabs():
    """the 1024x192x256 configuration on 8 slabs (23/24 planes each), seeded
    fields: the oracle run finishes in seconds on the host's cores"""
    g, s = make_case(C3, seed=3, P=8)
    return g, s


def test_slab_gs_real_shape_bitwise(tbl):
    g, s = _c3_slabs()
    rng = np.random.default_rng(11)
    Q = np.asfortranarray(s.Ut + 1e-2 * rng.standard_normal(s.Ut.shape))
    with gpu_from_case(g, s, helm_mode=tbl.TBL_HELM_LEX_SLAB, slabs=8) as sv:
        sv.set_field(tbl.F_Q, Q)
        sw, _ = sv.diffusion_solve(1e-30, 3)
        Ut = sv.get_field(tbl.F_UT)
    Uo = s.Ut.copy(order="F")
    swo, _ = g.diffusion_solve(Q, Uo, order=orc.GS_LEX_BLOCK, errorstop=1e-30, mmax=3)
    assert list(sw) == list(swo) == [3, 3, 3]
    assert ulps(Ut, Uo) == 0


def test_block_ssor_real_shape_bitwise(tbl):
    g, s = _c3_slabs()
    x = g.zeros()
    x[1:-1, 1:-1, 1:-1] = np.random.default_rng(12).standard_normal((g.nX - 2, g.nY - 2, g.nZ - 2))
    with gpu_from_case(g, s, precond=tbl.TBL_PC_SSOR_BLOCK, slabs=8) as sv:
        y = sv.apply_operator(1, x)
    yo = g.precond_apply(x, orc.PC_SSOR_BLOCK)
    I = (slice(1, -1),) * 3
    assert ulps(y[I], yo[I]) == 0
