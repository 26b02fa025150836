"""Pins of the oracle's stencil, boundary, diagnostic and recycling blocks.

Closed forms: polynomial exactness of each difference formula, telescoping of
the signed divergence sum, attenuation of a periodic mode by the edge average,
step / linear profiles for the integral thicknesses and friction velocities,
and the composed step's determinism and slab-count independence.
"""
import math

import numpy as np
import pytest

from oracle import orc
from tests.oracle_util import small_grid

I3 = (slice(1, -1),) * 3


def _field(g, fx=None, fy=None, fz=None, const=0.0):
    f = np.full((g.nX, g.nY, g.nZ), const, dtype=np.float64)
    if fx is not None:
        f = f + fx[:, None, None]
    if fy is not None:
        f = f + fy[None, :, None]
    if fz is not None:
        f = f + fz[None, None, :]
    return np.asfortranarray(f)


def _ut(*comps):
    return np.asfortranarray(np.stack(comps, axis=-1))


# ---------------------------------------------------------------- a3/a4
@pytest.mark.parametrize("U0", [0.004, 3.0])   # |Re U0 dx| below / above the Peclet switch
def test_convection_x_exact_on_quartic_plus_kk_dissipation(U0):
    g = small_grid(nX=16, nY=10, nZ=10)
    x = g.mesh["x1"]
    c = (0.3, -1.1, 0.7, 0.25, 0.4)            # phi = sum c_n x^n, degree 4
    phi = sum(cn * x ** n for n, cn in enumerate(c))
    dphi = sum(n * cn * x ** (n - 1) for n, cn in enumerate(c) if n)
    u = _field(g, const=U0)
    v = _field(g, fx=phi)                      # v depends on x only
    w = _field(g)
    conv, Q = g.momentum_rhs(_ut(u, v, w), g.zeros())
    got = conv[2:-2, 1:-1, 1:-1, 1]             # i = 3..nX-2 use the 5-point formulas
    Pe = abs(g.Re * U0 * g.dx)
    want = U0 * dphi[2:-2]
    if Pe > 2:   # Kawamura-Kuwahara dissipation |u| delta^4 phi / (4 dx) = |u| 24 c4 dx^3 / 4
        want = want + abs(U0) * 24 * c[4] * g.dx ** 3 / 4
    assert np.allclose(got, want[:, None, None], rtol=1e-9, atol=1e-9 * np.abs(want).max())
    # near the inflow/outflow (i = 2, nX-1) the central 2nd-order formula is exact on quadratics only
    assert conv[1, 3, 3, 1] == pytest.approx(U0 * (phi[2] - phi[0]) / (2 * g.dx), rel=1e-12)


def test_convection_z_uses_jacobian_and_span_periodic():
    g = small_grid(nX=10, nY=12, nZ=14)
    xi = g.mesh["Xi"]
    W0 = 0.002                                # Pe below 2 everywhere
    phi = 1.0 + 0.5 * xi - 0.2 * xi ** 2 + 0.1 * xi ** 3
    dphi = 0.5 - 0.4 * xi + 0.3 * xi ** 2
    u, v, w = _field(g), _field(g, fz=phi), _field(g, const=W0)
    conv, _ = g.momentum_rhs(_ut(u, v, w), g.zeros())
    want = W0 * dphi * g.mesh["JacInv"]
    assert np.allclose(conv[1:-1, 1:-1, 2:-2, 1], want[None, None, 2:-2], rtol=1e-9)
    # k = 2 and nZ-1 use the 2nd-order central formula
    for k in (1, g.nZ - 2):
        c2 = W0 * (phi[k + 1] - phi[k - 1]) / (2 * g.dXi) * g.mesh["JacInv"][k]
        assert conv[3, 3, k, 1] == pytest.approx(c2, rel=1e-12)
    # span: a periodic mode convected by a uniform v (4th-order central, Pe < 2)
    L = g.nY - 2
    jj = np.arange(g.nY)
    mode = np.sin(2 * math.pi * (jj - 1) / L)
    V0 = 0.001
    u, v, w = _field(g), _field(g, const=V0), _field(g, fy=mode)
    conv, _ = g.momentum_rhs(_ut(u, v, w), g.zeros())
    t = 2 * math.pi / L
    # modified wavenumber of the 4th-order (interior) and 2nd-order (j = 2, nY-1) central formulas
    k4 = (8 * math.sin(t) - math.sin(2 * t)) / (6 * g.dy)
    k2 = math.sin(t) / g.dy
    dmode = np.cos(2 * math.pi * (jj - 1) / L)
    assert np.allclose(conv[3, 2:-2, 3, 2], V0 * k4 * dmode[2:-2], atol=1e-12)
    assert conv[3, 1, 3, 2] == pytest.approx(V0 * k2 * dmode[1], abs=1e-12)
    assert conv[3, -2, 3, 2] == pytest.approx(V0 * k2 * dmode[-2], abs=1e-12)


def test_momentum_rhs_pressure_gradient_and_Q():
    g = small_grid()
    x, xi = g.mesh["x1"], g.mesh["Xi"]
    P = _field(g, fx=2.0 * x, fz=-3.0 * xi)
    rng = np.random.default_rng(1)
    Ut = np.asfortranarray(rng.uniform(-1e-3, 1e-3, (g.nX, g.nY, g.nZ, 3)))
    conv, Q = g.momentum_rhs(Ut, P)
    gp = [2.0, 0.0, None]
    assert np.allclose(Q[..., 0][I3], (Ut[..., 0] - g.dt * (conv[..., 0] + 2.0))[I3], rtol=1e-12)
    assert np.allclose(Q[..., 1][I3], (Ut[..., 1] - g.dt * conv[..., 1])[I3], rtol=1e-12)
    dz = -3.0 * g.mesh["JacInv"]
    assert np.allclose(Q[..., 2][I3], (Ut[..., 2] - g.dt * (conv[..., 2] + dz[None, None, :]))[I3], rtol=1e-12)


# ---------------------------------------------------------------- a6
def test_ppe_rhs_linear_fields():
    g = small_grid()
    x, x3 = g.mesh["x1"], g.mesh["x3"]
    Ut = _ut(_field(g, fx=3.0 * x), _field(g, const=1.0), _field(g, fz=-2.0 * x3))
    r = g.ppe_rhs(Ut)
    assert np.allclose(r[I3], (3.0 - 2.0) / g.dt, rtol=1e-9)
    assert np.all(r[0] == 0) and np.all(r[:, :, 0] == 0)


# ---------------------------------------------------------------- a10
def test_project_extrapolations_and_gradient():
    g = small_grid()
    x, xi = g.mesh["x1"], g.mesh["Xi"]
    # zero-slope quadratic at the wall / inflow, linear at the top / outflow
    PC = _field(g, fx=0.7 * (x - x[0]) ** 2, fz=1.5 * (xi - xi[0]) ** 2)
    PCin = PC.copy(order="F")
    Press = _field(g, const=2.0)
    Ut = np.asfortranarray(np.zeros((g.nX, g.nY, g.nZ, 3)))
    g.project(PC, Press, Ut)
    assert np.allclose(PC[1:-1, 1:-1, 0], PCin[1:-1, 1:-1, 0], rtol=1e-12, atol=1e-14)   # exact at k=1
    assert np.allclose(PC[0, 1:-1, 1:-1], PCin[0, 1:-1, 1:-1], rtol=1e-12, atol=1e-14)   # exact at i=1
    # i = nX: the 3-point formula is exact on linear data
    lin = _field(g, fx=0.3 * x + 1.0)
    g.project(lin, Press.copy(order="F"), Ut.copy(order="F"))
    assert np.allclose(lin[-1, 1:-1, 1:-1], 0.3 * x[-1] + 1.0, rtol=1e-12)
    # pressure update and central gradient on the interior
    assert np.allclose(Press[I3], 2.0 + PCin[I3], rtol=1e-14)
    # (i = nX-1 and k = nZ-1 read the outflow/top extrapolations, exact on linear data only)
    want_u = -g.dt * 0.7 * 2 * (x - x[0])
    assert np.allclose(Ut[1:-2, 1:-1, 1:-1, 0], want_u[1:-2, None, None], rtol=1e-10)
    want_w = -g.dt * 1.5 * 2 * (xi - xi[0]) * g.mesh["JacInv"]
    assert np.allclose(Ut[1:-1, 1:-1, 1:-2, 2], want_w[None, None, 1:-2], rtol=1e-10)
    assert np.allclose(Ut[1:-1, 1:-1, 1:-1, 1], 0.0)
    # span ghosts are copied before the inflow/outflow extrapolation (dnssetbl.f90:701-708),
    # so only i = 2..nX-1 of the ghost planes hold the periodic image
    assert np.array_equal(PC[1:-1, 0, :], PC[1:-1, -2, :]) and np.array_equal(PC[1:-1, -1, :], PC[1:-1, 1, :])


# ---------------------------------------------------------------- a11
def test_velocity_bcs():
    g = small_grid()
    rng = np.random.default_rng(4)
    Ut = np.asfortranarray(rng.standard_normal((g.nX, g.nY, g.nZ, 3)))
    ww = np.asfortranarray(rng.standard_normal((g.nX, g.nY)))
    g.apply_bcs(Ut, Ut.copy(order="F"), g.zeros(), wall_w=ww)
    assert np.all(Ut[:-1, 1:-1, 0, :2] == 0)
    # blowing/suction hook; the outflow BC (applied after the wall BC) overwrites i = nX
    assert np.array_equal(Ut[:-1, 1:-1, 0, 2], ww[:-1, 1:-1])
    top = (4.0 * Ut[:-1, 1:-1, -2, :] - Ut[:-1, 1:-1, -3, :]) / 3.0
    assert np.allclose(Ut[:-1, 1:-1, -1, :], top, rtol=1e-15)
    out = (4.0 * Ut[-2, 1:-1, :, :] - Ut[-3, 1:-1, :, :]) / 3.0
    assert np.allclose(Ut[-1, 1:-1, :, :], out, rtol=1e-15)
    assert np.array_equal(Ut[:, 0], Ut[:, -2]) and np.array_equal(Ut[:, -1], Ut[:, 1])


def test_pressure_bcs_closed_forms():
    g = small_grid(nX=12, nY=8, nZ=12)
    x, x3, xi = g.mesh["x1"], g.mesh["x3"], g.mesh["Xi"]
    zw = x3 - x3[0]
    # w = c (z - z_wall)^2 vanishes at the wall: bottom pressure = P(k=2) - (2c/Re) df1
    cW = 0.8
    u = _field(g, fx=0.5 + 0.01 * x)
    Ut = _ut(u, _field(g), _field(g, fz=cW * zw ** 2))
    Utold = Ut.copy(order="F")
    Press = _field(g, const=1.0)
    g.apply_bcs(Ut, Utold, Press)
    assert np.allclose(Press[1:-1, 1:-1, 0], 1.0 - (2 * cW / g.Re) * g.df1, rtol=1e-10)
    # top: 4-point one-sided d/dxi of the cubic through the 4 top nodes (after the top BC)
    wcol = Ut[3, 3, -4:, 2]
    dw = np.polyval(np.polyder(np.polyfit(xi[-4:], wcol, 3)), xi[-1])
    assert Press[3, 3, -1] == pytest.approx(dw * g.mesh["JacInv"][-1] / g.Re, rel=1e-7)
    # outflow: 3-point one-sided du/dx / Re
    ucol = Ut[-3:, 3, 4, 0]
    du = np.polyval(np.polyder(np.polyfit(x[-3:], ucol, 2)), x[-1])
    assert Press[-1, 3, 4] == pytest.approx(du / g.Re, rel=1e-8)
    # inflow: 2nd-order one-sided dP/dx = -(du/dt + u du/dx - visc) with u linear in x, v = w = 0
    Utold2 = Ut.copy(order="F")
    Utold2[0, :, :, 0] -= 0.02          # du/dt = 0.02/dt at the inflow
    P2 = _field(g, fx=0.1 * x ** 2)
    g.apply_bcs(Ut, Utold2, P2)
    k, j = 5, 3
    u1, ux = Ut[0, j, k, 0], 0.01
    dpdx = (0.02 / g.dt + u1 * ux)
    want = (2 * g.dx * dpdx - P2[2, j, k] + 4 * P2[1, j, k]) / 3.0
    assert P2[0, j, k] == pytest.approx(want, rel=1e-9)
    assert np.array_equal(Press[:, 0], Press[:, -2]) and np.array_equal(Press[:, -1], Press[:, 1])


# ---------------------------------------------------------------- a12
def test_divergence_sum_telescopes_to_boundary_values():
    g = small_grid(nX=14, nY=10, nZ=10)
    rng = np.random.default_rng(8)
    ux = rng.standard_normal(g.nX)
    Ut = _ut(_field(g, fx=ux), _field(g, const=0.7), _field(g))
    m, s = g.div_check(Ut)
    u = ux
    n = g.nX
    S = (8 * (u[n - 2] + u[n - 3] - u[2] - u[1]) - (u[n - 1] + u[n - 2] + u[n - 3] + u[n - 4] - u[0] - u[1] - u[2] - u[3])) / (12 * g.dx)
    S += (u[2] - u[0] + u[n - 1] - u[n - 3]) / (2 * g.dx)
    want = S * (g.nY - 2) * (g.nZ - 2)
    assert s == pytest.approx(want, rel=1e-10, abs=1e-9)
    assert m == pytest.approx(abs(want) / ((g.nX - 2) * (g.nY - 2) * (g.nZ - 2)), rel=1e-10)


# ---------------------------------------------------------------- a13
def test_mass_flux_uniform_stream():
    g = small_grid(nX=24)
    Ut = _ut(_field(g, const=2.5), _field(g), _field(g))
    m = g.mass_flux(Ut)
    x3, x2 = g.mesh["x3"], g.mesh["x2"]
    area = (g.nY - 2) * g.dy * (x3[-1] - x3[1])
    assert np.allclose(m, 2.5 * area / ((x3[-2] - x3[1]) * (x2[-2] - x2[1])), rtol=1e-12)


# ---------------------------------------------------------------- a14
def test_stats_edge_average_attenuates_mode():
    g = small_grid(nY=14)
    L = g.nY - 2
    jj = np.arange(g.nY)
    a, mm = 0.3, 2
    mode = a * np.cos(2 * math.pi * mm * (jj - 1) / L)
    U = np.linspace(0.0, 5.0, g.nZ)
    Ut = _ut(_field(g, fy=mode, fz=U), _field(g, const=-0.5), _field(g))
    sums = g.zeros2(9)
    g.stats_accumulate(Ut, sums)
    g.stats_accumulate(Ut, sums)
    att = math.cos(math.pi * mm / L)
    assert np.allclose(sums[..., 0], 2 * U[None, :], atol=1e-13)
    assert np.allclose(sums[..., 3], 2 * (U ** 2 + 0.5 * (a * att) ** 2)[None, :], rtol=1e-12)
    assert np.allclose(sums[..., 1], -1.0) and np.allclose(sums[..., 4], 0.5)
    assert np.allclose(sums[..., 6], 2 * (-0.5) * U[None, :], atol=1e-13)
    assert np.allclose(sums[..., 2], 0) and np.allclose(sums[..., 8], 0)


# ---------------------------------------------------------------- a15
def test_rescale_closed_forms():
    g = small_grid(nX=12, nY=8, nZ=16)
    x3 = g.mesh["x3"]
    nX, nZ, ir = g.nX, g.nZ, g.i_rec
    k0 = 9
    U = np.where(np.arange(nZ) >= k0, g.Ue, 0.0)
    sums = g.zeros2(9)
    sums[..., 0] = 3 * U[None, :]
    # near-wall slopes a (inflow) and b (recycling plane) give lambda = sqrt(a/b)
    a, b = 4.0, 9.0
    sums[0, 1, 0] = 3 * a * (x3[1] - x3[0])
    sums[ir - 1, 1, 0] = 3 * b * (x3[1] - x3[0])
    sums[0, 0, 0] = sums[ir - 1, 0, 0] = 0.0
    r = g.rescale(sums, 3, np.asfortranarray(np.zeros((nX, g.nY, nZ, 3))))
    assert r["lamda"] == pytest.approx(math.sqrt(a / b), rel=1e-12)
    assert np.all(r["bltkns"] == x3[k0])
    assert r["lamda03"] == 1.0
    # delta*: rectangle sum of (1 - U/Ue) dz up to the delta99 node; theta = 0 for a step
    i = 5
    assert r["deltastar"][i] == pytest.approx(x3[k0] - x3[0], rel=1e-12)
    assert r["theta"][i] == 0.0
    # Lund weight: 0 at the wall, 1/2 at eta = 0.2, 1 at eta = 1
    Wt = r["Wt"]
    assert Wt[i, 0] == pytest.approx(0.0, abs=1e-8)
    assert Wt[i, k0] == 1.0
    # lambda = lambda03 = 1 maps every k < nZ to itself
    sums2 = g.zeros2(9)
    prof = np.linspace(0.0, 1.2 * g.Ue, nZ)
    sums2[..., 0] = prof[None, :]
    r2 = g.rescale(sums2, 1, np.asfortranarray(np.zeros((nX, g.nY, nZ, 3))))
    assert r2["lamda"] == pytest.approx(1.0) and r2["lamda03"] == pytest.approx(1.0)
    assert list(r2["newk"][: nZ - 1]) == list(range(1, nZ)) and r2["newk"][nZ - 1] == nZ - 1
    assert np.allclose(r2["avg"][..., 0], prof[None, :])


def test_lund_weight_special_points():
    from turbulentbl_amd.inflow import lund_weight
    assert lund_weight(0.2) == pytest.approx(0.5, abs=1e-15)
    assert lund_weight(1.0) == pytest.approx(1.0, abs=1e-15)
    assert lund_weight(0.0) == pytest.approx(0.0, abs=1e-15)


# ---------------------------------------------------------------- a16
def test_ke_and_guard():
    g = small_grid()
    avg = g.zeros2(9)
    avg[..., 0] = 7.0
    Ut = _ut(_field(g, const=7.0 + 0.5), _field(g), _field(g))
    uf = np.zeros((g.nY, g.nZ, 3), order="F")
    ke, gd = g.ke_guard(Ut, avg, uf)
    assert ke == pytest.approx(0.25 * (g.nX - 2) / g.nX, rel=1e-14)
    assert gd == 0
    assert np.allclose(uf[1:-1, 1:-1, 0], 0.5)
    Ut[4, 3, 3, 1] = 26.0
    assert g.ke_guard(Ut, avg)[1] == 1
    Ut[4, 3, 3, 1] = -26.0           # the reference's guard is on the signed maximum
    assert g.ke_guard(Ut, avg)[1] == 0


# ---------------------------------------------------------------- a17
def test_wall_shear_exact_on_quadratics():
    g = small_grid(nX=16)
    x, x3 = g.mesh["x1"], g.mesh["x3"]
    zw = x3 - x3[0]
    a, b = 2.0, -0.7
    u = _field(g, fz=a * zw + b * zw ** 2)
    w = g.zeros()
    w[:, :, 0] = (0.3 * x ** 3)[:, None]
    tau = g.wall_shear(_ut(u, g.zeros(), w))
    dw = 0.9 * x ** 2
    ny = g.nY - 2
    mid = slice(3, g.nX - 3)          # 6th-order interior formula: exact on the cubic
    assert np.allclose(tau[mid], ny * (a + dw[mid]) / g.Re, rtol=1e-9)
    i = 1                             # 2nd-order near the ends: cubic error dx^2 * w'''/6
    want = ny * (a + dw[i] + 0.3 * g.dx ** 2) / g.Re
    assert tau[i] == pytest.approx(want, rel=1e-9)


# ---------------------------------------------------------------- a2
def test_recycle_inflow_hand_values():
    g = small_grid(nX=12, nY=8, nZ=10, freq_tim_avg=5)
    nZ, ir = g.nZ, g.i_rec
    avg = g.zeros2(9)
    avg[ir - 1, :, 0] = np.linspace(1.0, 10.0, nZ)     # all below 0.99 Ue
    avg[ir - 1, :, 1] = 0.25
    avg[ir - 1, :, 2] = -0.125
    avg[ir - 1, 7, 0] = 0.995 * g.Ue                   # k = 8 takes the "outer" branch
    Wt = g.zeros2()
    Wt[0, :] = 0.3
    newk = np.arange(1, nZ + 1, dtype=np.int32)
    newk[3] = 6
    Ume = np.full(nZ, 2.0)
    uf = np.zeros((g.nY, nZ, 3), order="F")
    uf[:, 5, 0] = 0.5                                   # fluctuation on plane new(4) = 6
    Ut = np.asfortranarray(np.zeros((g.nX, g.nY, nZ, 3)))
    g.recycle_inflow(Ut, 6, 2.0, avg, Wt, newk, Ume, uf)    # nstep 6 > freq 5, eps = 3e-4
    eps = 6 * 5e-5
    k = 4
    um = eps * avg[ir - 1, 5, 0] + (1 - eps) * 2.0
    assert np.allclose(Ut[0, 1:-1, k - 1, 0], um + 2.0 * 0.5, rtol=1e-14)
    assert np.allclose(Ut[0, 1:-1, k - 1, 1], 2.0 * 0.25, rtol=1e-14)
    assert np.allclose(Ut[0, 1:-1, k - 1, 2], -0.125, rtol=1e-14)
    assert np.allclose(Ut[0, 1:-1, 7, 0], 0.995 * g.Ue)    # outer: plain copy of the means
    assert np.all(Ut[0, 1:-1, 0] == 0) and np.all(Ut[0, 1:-1, -1] == 0)
    Ut2 = np.asfortranarray(np.zeros_like(Ut))
    g.recycle_inflow(Ut2, 5, 2.0, avg, Wt, newk, Ume, uf)   # nstep <= freq: untouched
    assert not Ut2.any()


# ------------------------------------------------- a3/a4, dnstbl variant
# dnstbl/dnstbl.f90:620-766: the Kawamura-Kuwahara branch only inside the
# cell-Peclet window, Pe signed by sign(1, ceiling(u)).  On a quartic the
# 4th-order central formula is exact and the Kawamura-Kuwahara one adds
# |u| 24 c4 dx^3 / 4, so the exact derivative tells which branch ran.
def _x_quartic_case(g, U0):
    x = g.mesh["x1"]
    c = (0.3, -1.1, 0.7, 0.25, 0.4)
    phi = sum(cn * x ** n for n, cn in enumerate(c))
    dphi = sum(n * cn * x ** (n - 1) for n, cn in enumerate(c) if n)
    conv, _ = g.momentum_rhs(_ut(_field(g, const=U0), _field(g, fx=phi), _field(g)), g.zeros())
    got = conv[2:-2, 1:-1, 1:-1, 1]
    exact = U0 * dphi[2:-2]
    kk = exact + abs(U0) * 24 * c[4] * g.dx ** 3 / 4
    return got, exact[:, None, None], kk[:, None, None], conv, phi


@pytest.mark.parametrize("pe,branch", [(0.5, "central"), (10.0, "kk"), (49.0, "kk"), (60.0, "central"),
                                       (-10.0, "kk"), (-60.0, "central")])
def test_dnstbl_peclet_window_x(pe, branch):
    g = small_grid(nX=16, nY=10, nZ=10, scheme=1, Re=1.0)
    U0 = pe / (g.Re * g.dx)                  # |U0| >= 1 for |pe| >= 10 (Re dx = 0.8): sgn = sign(u)
    assert abs(U0) >= 1.0 or pe > 0
    got, exact, kk, _, _ = _x_quartic_case(g, U0)
    want = kk if branch == "kk" else exact
    assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_dnstbl_peclet_sign_quirk_for_small_negative_u():
    """-1 < u < 0: ceiling(u) = 0, sign(1, 0) = +1, so Pe = Re u dx < 0 and the
    central formula runs even where |Re u dx| > 2 (dnssetbl takes KK there)"""
    g1 = small_grid(nX=16, nY=10, nZ=10, scheme=1)
    g0 = small_grid(nX=16, nY=10, nZ=10)
    U0 = -0.5
    assert abs(g1.Re * U0 * g1.dx) > 2.0
    got1, exact, kk, _, _ = _x_quartic_case(g1, U0)
    got0, _, _, _, _ = _x_quartic_case(g0, U0)
    assert np.allclose(got1, exact, rtol=1e-9, atol=1e-9 * np.abs(exact).max())
    assert np.allclose(got0, kk, rtol=1e-9, atol=1e-9 * np.abs(kk).max())


def test_dnstbl_wall_normal_window_is_2_to_5():
    g = small_grid(nX=10, nY=10, nZ=16, scheme=1)
    xi, x3 = g.mesh["Xi"], g.mesh["x3"]
    c = (0.2, 0.5, -0.3, 0.1, 0.05)
    phi = sum(cn * xi ** n for n, cn in enumerate(c))
    dphi = sum(n * cn * xi ** (n - 1) for n, cn in enumerate(c) if n)
    dz = x3[1:] - x3[:-1]
    for pe_target, kk_expected in ((3.5, True), (7.0, False)):
        W0 = pe_target / (g.Re * dz[7])           # Pe at k = 8 (1-based) is pe_target
        conv, _ = g.momentum_rhs(_ut(_field(g), _field(g, fz=phi), _field(g, const=W0)), g.zeros())
        k = 7
        pe = g.Re * W0 * dz[k]
        exact = W0 * dphi[k] * g.mesh["JacInv"][k]
        kk = (W0 * dphi[k] + abs(W0) * 24 * c[4] * g.dXi ** 3 / 4) * g.mesh["JacInv"][k]
        assert (2.0 < pe < 5.0) == kk_expected
        assert conv[4, 4, k, 1] == pytest.approx(kk if kk_expected else exact, rel=1e-9)


def test_dnstbl_span_edges_first_order_upwind_and_streamwise_edges():
    """j = 2, nY-1: a1 (F(j) - F(j-1)) / dy + b (F(j+1) - F(j)) / dy with
    a1 = max(v, 0), b = min(v, 0), exact on a field linear in y away from the
    periodic seam; i = 2, nX-1: u (F(i+1) - F(i-1)) / (2 dx)"""
    g = small_grid(nX=12, nY=12, nZ=10, scheme=1)
    y = g.mesh["x2"]
    phi = 0.7 + 2.5 * y
    for V0, j in ((0.003, g.nY - 2), (-0.003, 1)):   # backward at j = nY-1, forward at j = 2 (0-based)
        conv, _ = g.momentum_rhs(_ut(_field(g), _field(g, const=V0), _field(g, fy=phi)), g.zeros())
        assert conv[4, j, 4, 2] == pytest.approx(V0 * 2.5, rel=1e-10)
    x = g.mesh["x1"]
    phi = 1.0 + 0.5 * x - 0.25 * x ** 2
    U0 = 0.002
    conv, _ = g.momentum_rhs(_ut(_field(g, const=U0), _field(g, fx=phi), _field(g)), g.zeros())
    for i in (1, g.nX - 2):
        assert conv[i, 4, 4, 1] == pytest.approx(U0 * (phi[i + 1] - phi[i - 1]) / (2 * g.dx), rel=1e-12)


def test_dnstbl_body_force_and_Q_order():
    """Q = (U + dt (forc - conv)) - dt gradp: the force enters as dt forc per
    component, and without it the variant's Q equals U - dt (conv + gradp) to
    rounding"""
    rng = np.random.default_rng(4)
    g0 = small_grid(scheme=1)
    forc = (0.3, -0.2, 0.05)
    g1 = small_grid(scheme=1, forc=forc)
    x = g0.mesh["x1"]
    P = _field(g0, fx=2.0 * x)
    Ut = np.asfortranarray(rng.uniform(-1e-3, 1e-3, (g0.nX, g0.nY, g0.nZ, 3)))
    conv0, Q0 = g0.momentum_rhs(Ut, P)
    conv1, Q1 = g1.momentum_rhs(Ut, P)
    assert np.array_equal(conv0, conv1)
    for p in range(3):
        assert np.allclose((Q1 - Q0)[..., p][I3], g0.dt * forc[p], rtol=1e-6)
    gp = 2.0
    assert np.allclose(Q0[..., 0][I3], (Ut[..., 0] - g0.dt * (conv0[..., 0] + gp))[I3], rtol=1e-12)


def test_wall_shear_window_mean_and_cf():
    """The oracle's wall-shear window (dnstbl.f90:1666-1682): at each step with
    nstep % window == 0 it holds the window's time mean of tau(i) per span row
    and C_f = 2 tau_w / Ue^2 (tau_w in units of rho u_tau^2, Ue = U_inf/u_tau),
    checked against tau re-evaluated on each step's field; between windows
    the sum restarts from zero."""
    from tests.case_util import make_case
    g, s = make_case((20, 10, 14), seed=2, nstep0=10)
    s.tau_window = 3
    taus, ready, since = [], [], []
    for _ in range(7):                      # steps 11..17: windows end at 12 and 15
        d = s.step()
        taus.append(g.wall_shear(s.Ut))
        since.append(taus[-1])
        ready.append(d["cf_ready"])
        if d["cf_ready"]:
            assert s.nstep % 3 == 0
            # the reference divides by the window length even when the run
            # started inside a window (the first window here holds 2 steps)
            want = np.sum(since, axis=0) / 3.0 / (g.nY - 2)
            assert np.allclose(s.taumean, want, rtol=1e-14, atol=0)
            assert np.allclose(s.cf, 2.0 * want / g.Ue ** 2, rtol=1e-14, atol=0)
            since = []
    assert ready == [False, True, False, False, True, False, False]
    assert s.nwin == 2
    # two steps into the third window the sum holds exactly those two steps
    assert np.allclose(s.tausum[1:-1], (taus[-2] + taus[-1])[1:-1], rtol=1e-14)
    assert s.taumean[0] == 0.0 and s.cf[-1] == 0.0
