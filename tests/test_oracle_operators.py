"""Pins of the CPU oracle's mesh, stencils and linear solvers.

The reference holds no tests or fixtures, and it cannot be built here (both
solvers `use mpi`, which amdflang would need a stand-in module for, and
dnssetbl.f90:98,101 uses the Intel-only INQUIRE(directory=)).  Each check
below pins an oracle function against something the mathematics fixes:
closed forms, polynomial exactness, dense linear algebra on tiny grids.
"""
import math

import numpy as np
import pytest

from oracle import orc
from tests.oracle_util import (dense_operator, from_interior, ijk_of, interior_index,
                               small_grid, to_interior)


# ---------------------------------------------------------------- a1: mesh
def test_mesh_uniform_and_tanh_map_derivatives():
    g = orc.Grid(20, 12, 2001)
    m = g.mesh
    assert np.allclose(np.diff(m["x1"]), 12.0 / 19, rtol=0, atol=1e-13)
    assert np.allclose(np.diff(m["x2"]), 4.0 / 11, rtol=0, atol=1e-13)
    dxi = g.dXi
    x3 = m["x3"]
    # Jac = dx3/dxi and Jac2 = d2x3/dxi2 of the map (second-order FD of the nodes)
    fd1 = (x3[2:] - x3[:-2]) / (2 * dxi)
    fd2 = (x3[2:] - 2 * x3[1:-1] + x3[:-2]) / dxi ** 2
    assert np.allclose(m["Jac"][1:-1], fd1, rtol=2e-6, atol=0)
    assert np.allclose(m["Jac2"][1:-1], fd2, rtol=0, atol=2e-5 * np.abs(fd2).max())
    assert np.all(np.diff(x3) > 0)
    assert abs(x3[-1] - 3.4) < 1e-6
    # the REAL(4) literal 3.4 in (Xi-3.4) shifts the wall node by -J(0)*(3.4f-3.4)
    shift = float(np.float32(3.4)) - 3.4
    assert x3[0] == pytest.approx(-m["Jac"][0] * shift, rel=1e-3)
    assert np.allclose(m["JacInv"], 1.0 / m["Jac"], rtol=1e-15)


def test_coefficient_identities():
    g = orc.Grid(16, 10, 24)
    m = g.mesh
    J, J2 = m["Jac"], m["Jac2"]
    assert np.allclose(m["Ab"] + m["At"], 2.0 / (g.dXi ** 2 * J ** 2), rtol=1e-14)
    assert np.allclose(m["Ab"] - m["At"], J2 / (g.dXi * J ** 3), rtol=1e-12, atol=1e-9)
    assert np.allclose(m["Ap"], -(g.Aw + g.Ae + g.As + g.An + m["Ab"] + m["At"]), rtol=1e-14)
    assert g.Aw == g.Ae == 1.0 / g.dx ** 2 and g.As == g.An == 1.0 / g.dy ** 2
    assert np.allclose(m["rAp"] * m["Ap"], 1.0, rtol=1e-15)


# ---------------------------------------------------------------- a7: Poisson operator
def test_poisson_apply_polynomial_and_fourier_exactness():
    g = orc.Grid(12, 14, 16)
    m = g.mesh
    x1, x2, xi = m["x1"], m["x2"], m["Xi"]
    J, J2 = m["Jac"], m["Jac2"]
    I = (slice(1, -1),) * 3
    # x^2 -> 2 exactly (uniform, Dirichlet neighbours are just values)
    phi = np.asfortranarray(np.broadcast_to(x1[:, None, None] ** 2, (g.nX, g.nY, g.nZ)).copy())
    assert np.allclose(g.poisson_apply(phi)[I], 2.0, rtol=1e-9)
    # xi^2 -> 2/J^2 - 2 xi J''/J^3 (chain rule on the stretched map, exact for quadratics in xi)
    phi = np.asfortranarray(np.broadcast_to(xi[None, None, :] ** 2, (g.nX, g.nY, g.nZ)).copy())
    want = 2 / J ** 2 - 2 * xi * J2 / J ** 3
    got = g.poisson_apply(phi)[I]
    assert np.allclose(got, np.broadcast_to(want[None, None, 1:-1], got.shape), rtol=1e-9)
    # periodic span mode: eigenvalue (2cos(2 pi m/(nY-2)) - 2)/dy^2
    L = g.nY - 2
    for mm in (1, 3):
        jj = np.arange(g.nY)
        mode = np.cos(2 * math.pi * mm * (jj - 1) / L)
        phi = np.asfortranarray(np.broadcast_to(mode[None, :, None], (g.nX, g.nY, g.nZ)).copy())
        lam = (2 * math.cos(2 * math.pi * mm / L) - 2) / g.dy ** 2
        got = g.poisson_apply(phi)[I]
        assert np.allclose(got, lam * phi[I], atol=1e-9 * abs(lam))


# ---------------------------------------------------------------- a5: Helmholtz (GS)
def _helm_dense(g):
    A, idx = dense_operator(g, g.poisson_apply)
    H = np.eye(len(idx)) - (g.dt / g.Re) * A
    return H, A, idx


def _random_fields(g, seed=0):
    rng = np.random.default_rng(seed)
    Ut = np.asfortranarray(rng.standard_normal((g.nX, g.nY, g.nZ, 3)))
    Q = np.asfortranarray(rng.standard_normal((g.nX, g.nY, g.nZ, 3)))
    return Ut, Q


def _boundary_rhs(g, A_apply, u, idx):
    """Contribution of the fixed (non-interior) values of u to A u on the interior."""
    ub = u.copy(order="F")
    ub.reshape(-1, order="F")[idx] = 0.0
    return to_interior(g, A_apply(ub), idx)


@pytest.mark.parametrize("order", [orc.GS_LEX_BLOCK, orc.GS_RB, orc.GS_ZLINE_RB, orc.GS_ZEBRA])
def test_gs_converges_to_dense_helmholtz_solution(order):
    g = small_grid(dt=2e-3)   # larger dt/Re so the off-diagonal part matters
    H, A, idx = _helm_dense(g)
    Ut, Q = _random_fields(g)
    Ut0 = Ut.copy(order="F")
    sweeps, res = g.diffusion_solve(Q, Ut, order=order, errorstop=1e-15, mmax=400)
    for p in range(3):
        b = to_interior(g, Q[..., p], idx) + (g.dt / g.Re) * _boundary_rhs(g, g.poisson_apply, Ut0[..., p], idx)
        u = np.linalg.solve(H, b)
        assert np.allclose(to_interior(g, Ut[..., p], idx), u, rtol=0, atol=1e-11 * np.abs(u).max())
        # boundary values untouched
        assert np.array_equal(Ut[0, :, :, p], Ut0[0, :, :, p])
        assert np.array_equal(Ut[:, :, 0, p], Ut0[:, :, 0, p])


def _sweep_split(g, H, idx, order, P):
    """Lower / diagonal / upper parts of H as one in-place sweep sees them."""
    n = len(idx)
    I, J, K = ijk_of(g, idx)
    if order == orc.GS_RB:
        colour = (I + J + K) % 2
        rank = colour * n + np.arange(n)         # red (even) first, then black
    elif order == orc.GS_ZLINE_RB:
        colour = (I + J) % 2                     # columns: red (even i+j) first
        rank = colour * n + np.arange(n)
    else:
        rank = np.arange(n)
    slab = np.zeros(n, dtype=int)
    if order == orc.GS_LEX_BLOCK:
        for r in range(P):
            slab[(K >= g.g_bn[r]) & (K <= g.g_en[r])] = r
    before = rank[None, :] < rank[:, None]
    same = slab[None, :] == slab[:, None]
    if order == orc.GS_ZEBRA:
        # span-row zebra: columns solved exactly (block diagonal); a column of
        # the second pass (odd 1-based j) sees the first pass's new values
        # (its span neighbours); x neighbours, in the same pass, stay old
        col = (I[:, None] == I[None, :]) & (J[:, None] == J[None, :])
        D = np.where(col, H, 0.0)
        newer = ((J % 2)[:, None] == 1) & ((J % 2)[None, :] == 0)
        L = np.where(newer & ~col, H, 0.0)
        U = H - L - D
        return L, D, U
    if order == orc.GS_ZLINE_RB:
        # block Gauss-Seidel: a column's own couplings are solved exactly (block diagonal)
        col = (I[:, None] == I[None, :]) & (J[:, None] == J[None, :])
        D = np.where(col, H, 0.0)
        L = np.where(before & ~col, H, 0.0)
        U = H - L - D
        return L, D, U
    L = np.where(before & same, H, 0.0)
    D = np.diag(np.diag(H))
    U = H - L - D
    return L, D, U


@pytest.mark.parametrize("order,P", [(orc.GS_LEX_BLOCK, 1), (orc.GS_LEX_BLOCK, 2), (orc.GS_RB, 1),
                                     (orc.GS_ZLINE_RB, 1), (orc.GS_ZEBRA, 1)])
def test_single_gs_sweep_is_the_matrix_splitting(order, P):
    g = small_grid(nZ=12, P=P, dt=2e-3)
    # slab bounds as the reference computes them (dnssetbl.f90:76-92)
    it1, it2 = g.nZ // P, P - g.nZ % P
    b, e = [1], []
    for r in range(P):
        if r == it2:
            it1 += 1
        e.append(b[r] + it1 - 1)
        b.append(e[r] + 1)
    g.g_bn = [b[0] + 1] + b[1:P]
    g.g_en = e[:P - 1] + [e[P - 1] - 1]
    H, A, idx = _helm_dense(g)
    Ut, Q = _random_fields(g, 3)
    Ut0 = Ut.copy(order="F")
    g.helm_sweep(Q, Ut, 2, order)
    b_ = to_interior(g, Q[..., 1], idx) + (g.dt / g.Re) * _boundary_rhs(g, g.poisson_apply, Ut0[..., 1], idx)
    L, D, U = _sweep_split(g, H, idx, order, P)
    want = np.linalg.solve(L + D, b_ - U @ to_interior(g, Ut0[..., 1], idx))
    assert np.allclose(to_interior(g, Ut[..., 1], idx), want, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("order", [orc.GS_ZLINE_RB, orc.GS_ZEBRA])
@pytest.mark.parametrize("nZ", [6, 19, 40, 56])
def test_zline_sweep_partitioned_solve_is_block_gs(nZ, order):
    """The z-line sweep splits each column into chunks (nZ-2 = 4: one chunk;
    17: two chunks of 8/9; 38: three; 54: four) and still performs one exact
    block Gauss-Seidel step with the columns as blocks."""
    g = small_grid(nX=7, nY=6, nZ=nZ, dt=2e-3)
    H, A, idx = _helm_dense(g)
    Ut, Q = _random_fields(g, 11)
    Ut0 = Ut.copy(order="F")
    g.helm_sweep(Q, Ut, 3, order)
    b_ = to_interior(g, Q[..., 2], idx) + (g.dt / g.Re) * _boundary_rhs(g, g.poisson_apply, Ut0[..., 2], idx)
    L, D, U = _sweep_split(g, H, idx, order, 1)
    want = np.linalg.solve(L + D, b_ - U @ to_interior(g, Ut0[..., 2], idx))
    assert np.allclose(to_interior(g, Ut[..., 2], idx), want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_zline_partition_rule():
    assert [orc.zline_chunks(n) for n in (2, 16, 17, 190, 256, 257, 512)] == [1, 1, 2, 12, 16, 9, 16]
    # balanced chunks of at least 2 rows covering 2..nZ-1
    for nz in (2, 3, 17, 33, 190, 254, 300, 512):
        C = orc.zline_chunks(nz)
        b = [orc.zline_begin(c, C, nz) for c in range(C + 1)]
        assert b[0] == 2 and b[-1] == nz + 2
        sizes = np.diff(b)
        assert sizes.min() >= 2 and sizes.max() - sizes.min() <= 1 and sizes.max() <= (16 if nz <= 256 else 32)


def test_helm_residual_is_l1_of_dense_residual():
    g = small_grid(dt=2e-3)
    H, A, idx = _helm_dense(g)
    Ut, Q = _random_fields(g, 5)
    for p in (1, 3):
        b = to_interior(g, Q[..., p - 1], idx) + (g.dt / g.Re) * _boundary_rhs(g, g.poisson_apply, Ut[..., p - 1], idx)
        r = b - H @ to_interior(g, Ut[..., p - 1], idx)
        assert g.helm_residual(Q, Ut, p) == pytest.approx(np.abs(r).sum(), rel=1e-12)


# ---------------------------------------------------------------- a8: preconditioners
def _ssor_dense(g, A, idx, order, P, omega):
    n = len(idx)
    I, J, K = ijk_of(g, idx)
    Ablk = A.copy()
    if order == "lex":
        # per-slab, non-periodic span: drop the wrap couplings and cross-slab couplings
        wrap = np.abs(J[:, None] - J[None, :]) > 1
        slab = np.zeros(n, dtype=int)
        for r in range(P):
            slab[(K >= g.g_bn[r]) & (K <= g.g_en[r])] = r
        cross = slab[:, None] != slab[None, :]
        Ablk[wrap | cross] = 0.0
        rank = np.arange(n)
    else:
        rank = ((I + J + K) % 2) * n + np.arange(n)
    before = rank[None, :] < rank[:, None]
    D = np.diag(np.diag(Ablk))
    L = np.where(before, Ablk, 0.0) - np.where(before, D, 0.0)
    U = Ablk - L - D
    return np.linalg.solve(D / omega + U, ((2 - omega) / omega) * D @ np.linalg.inv(D / omega + L))


@pytest.mark.parametrize("pc,P", [(orc.PC_SSOR_BLOCK, 1), (orc.PC_SSOR_BLOCK, 2), (orc.PC_SSOR_RB, 1)])
def test_ssor_equals_product_form(pc, P):
    g = small_grid(nZ=12, P=P)
    it1, it2 = g.nZ // P, P - g.nZ % P
    b, e = [1], []
    for r in range(P):
        if r == it2:
            it1 += 1
        e.append(b[r] + it1 - 1)
        b.append(e[r] + 1)
    g.g_bn = [b[0] + 1] + b[1:P]
    g.g_en = e[:P - 1] + [e[P - 1] - 1]
    A, idx = dense_operator(g, g.poisson_apply)
    Minv = _ssor_dense(g, A, idx, "lex" if pc == orc.PC_SSOR_BLOCK else "rb", P, g.omega)
    rng = np.random.default_rng(11)
    for _ in range(2):
        r = from_interior(g, rng.standard_normal(len(idx)), idx)
        got = to_interior(g, g.precond_apply(r, pc), idx)
        want = Minv @ to_interior(g, r, idx)
        assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_ssor_omega_is_real4_literal():
    assert orc.OMEGA_REF == 1.83500003814697265625


def test_jacobi_and_none():
    g = small_grid()
    rng = np.random.default_rng(2)
    idx = interior_index(g)
    r = from_interior(g, rng.standard_normal(len(idx)), idx)
    assert np.array_equal(g.precond_apply(r, orc.PC_NONE), r)
    got = g.precond_apply(r, orc.PC_JACOBI)
    assert np.allclose(got[1:-1, 1:-1, 1:-1], r[1:-1, 1:-1, 1:-1] / g.mesh["Ap"][None, None, 1:-1], rtol=1e-15)


# ---------------------------------------------------------------- a9: GMRES
def _poisson_problem(g, seed):
    A, idx = dense_operator(g, g.poisson_apply)
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(len(idx)) * 1e3
    return A, idx, b


@pytest.mark.parametrize("pc", [orc.PC_SSOR_BLOCK, orc.PC_SSOR_RB, orc.PC_NONE])
def test_gmres_tight_solve_matches_dense_solution(pc):
    g = small_grid()
    A, idx, b = _poisson_problem(g, 7)
    x = np.linalg.solve(A, b)
    pRhs = from_interior(g, b, idx)
    PC = from_interior(g, np.random.default_rng(1).standard_normal(len(idx)), idx)   # warm start
    PC[0, :, :] = 5.0   # faces must be zeroed by the solver (dnssetbl.f90:451-460)
    st = g.ppe_gmres(pRhs, PC, pc=pc, errtol=1e-13, kmax=20, mmax=200)
    assert np.allclose(to_interior(g, PC, idx), x, rtol=0, atol=1e-9 * np.abs(x).max())
    assert np.all(PC[0] == 0) and np.all(PC[-1] == 0) and np.all(PC[:, :, 0] == 0)
    assert st["true_res"] <= 1e-9 * np.linalg.norm(b)


@pytest.mark.parametrize("kit", [1, 3, 6])
def test_one_gmres_cycle_is_the_krylov_least_squares_minimiser(kit):
    """After one cycle of kit steps from x0=0, x minimises ||M(b - A x)|| over K_kit(MA, Mb)."""
    g = small_grid()
    A, idx, b = _poisson_problem(g, 9)
    Minv, _ = dense_operator(g, lambda r: g.precond_apply(r, orc.PC_SSOR_BLOCK))
    pRhs = from_interior(g, b, idx)
    PC = g.zeros()
    g.ppe_gmres(pRhs, PC, pc=orc.PC_SSOR_BLOCK, errtol=1e-300, kmax=kit, mmax=1)
    MA, Mb = Minv @ A, Minv @ b
    K = np.zeros((len(b), kit))
    v = Mb / np.linalg.norm(Mb)
    for q in range(kit):   # orthonormal Krylov basis by QR (independent of Arnoldi/Givens)
        K[:, q] = v
        Qm, _ = np.linalg.qr(K[:, :q + 1])
        v = MA @ Qm[:, q]
    Qm, _ = np.linalg.qr(K)
    y, *_ = np.linalg.lstsq(MA @ Qm, Mb, rcond=None)
    want = Qm @ y
    assert np.allclose(to_interior(g, PC, idx), want, rtol=0, atol=1e-8 * np.abs(want).max())


def test_gmres_stopping_point_is_first_krylov_dimension_below_tolerance():
    """x0 = 0, one cycle: the Givens estimate |g(k+1)| is min over K_k of ||M(b - A x)||, so the
    inner loop stops at the first k whose least-squares residual / ||M b|| <= errtol."""
    g = small_grid(nX=14, nY=10, nZ=12)
    A, idx, b = _poisson_problem(g, 13)
    Minv, _ = dense_operator(g, lambda r: g.precond_apply(r, orc.PC_SSOR_BLOCK))
    MA, Mb = Minv @ A, Minv @ b
    kmax, tol = 30, 1e-6
    K = np.zeros((len(b), kmax))
    v = Mb / np.linalg.norm(Mb)
    first = None
    for q in range(kmax):
        K[:, q] = v
        Qm, _ = np.linalg.qr(K[:, :q + 1])
        y, *_ = np.linalg.lstsq(MA @ Qm, Mb, rcond=None)
        if first is None and np.linalg.norm(Mb - MA @ Qm @ y) / np.linalg.norm(Mb) <= tol:
            first = q + 1
        v = MA @ Qm[:, q]
    assert first is not None
    PC = g.zeros()
    st = g.ppe_gmres(from_interior(g, b, idx), PC, pc=orc.PC_SSOR_BLOCK, errtol=tol, kmax=kmax, mmax=1)
    assert st["iters"] == first and st["cycles"] == 1
    assert st["sumav"] <= tol
