"""The five benchmark configurations with their physical set-up.

Axes.  A configuration written A x B x C (streamwise x wall-normal x span)
maps to the reference's (nX, nY, nZ) = (A, C, B): j is the periodic span, k
the wall-normal direction (dnssetbl/dnssetbl.f90:183-243).

Units (the reference's, dnssetbl.f90:28-30): lengths in inflow boundary-layer
thicknesses delta99, velocities in inflow friction velocities u_tau, so that
Re = u_tau delta99 / nu (= Re_tau) and Ue = U_inf / u_tau; the momentum
thickness Reynolds number is theta * Re * Ue (:1151-1153).

Mapping Re_theta,in -> (Re, Ue):

* turbulent inflow (log-law mean, configs 2-5): Ue = U_inf/u_tau from the
  Coles-Fernholz skin-friction relation U_inf/u_tau = ln(Re_theta)/0.384 +
  4.127 (Nagib, Chauhan & Monkewitz 2007); Re is then the root of
  theta(Re) * Re * Ue = Re_theta, theta the momentum thickness of the build's
  own mean profile (inflow.mean_profile: Reichardt inner law + Coles wake
  reaching Ue at delta99), so that the inflow measured by the reference's
  integral carries the requested Re_theta;
* Blasius inflow (config 1): with delta99 = 1 the Blasius similarity
  variable is eta = eta99 * z (eta99 = 4.910); theta = 0.6641/eta99 and
  Re * Ue = U_inf delta99 / nu = Re_theta / theta; the wall shear
  tau_w = mu U_inf f''(0) eta99 / delta99 gives Ue^2 = Re*Ue / (f''(0) eta99).

Domain 60 x 3 x 6 delta99 (streamwise, wall-normal, span): Lx = 60, Lz = 3
(wall-normal), Ly = 6 (span).  The wall-normal map is the reference's
x3 = L [1 + tanh(delta (xi - L)/2) / tanh(delta L/2)], xi in [0, L], i.e.
x3 = L [1 - tanh(g (1 - eta)) / tanh g] with eta = xi/L and stretch factor
g = delta L / 2 = 2.5 (delta = 5/3); its first cell is at 1.1e-3 delta99 at
nZ = 192 (y+ below 1 for every configuration, reported by ``describe``).

Time step: fixed for the run from the initial field, CFL 0.3 in the
multi-dimensional convective form dt * max(|u|/dx + |v|/dy + |w|/dz_k) = 0.3
over the interior nodes, dz_k = (x3(k+1) - x3(k-1))/2 (the reference keeps
dt fixed; dnssetbl.f90:27).

Recycling plane at 0.8 Lx: i_rec = round(0.8 (nX - 1)) + 1.
"""
import math

import numpy as np

from . import inflow

DOMAIN = dict(Lx=60.0, Ly=6.0, Lz=3.0)      # streamwise, span, wall-normal (delta99)
STRETCH = 2.5                               # g = delta * Lz / 2
CFL = 0.3

# Blasius constants (shooting solution of f''' + f f''/2 = 0): f''(0), eta at
# f' = 0.99, momentum thickness in eta units
BLASIUS_FPP0 = 0.332057
BLASIUS_ETA99 = 4.9100
BLASIUS_THETA_ETA = 0.66411

# (nX, nY span, nZ wall-normal)
CASES = {
    "c1": dict(grid=(64, 32, 48), name="64x48x32", re_theta=300.0, ic="blasius", noise=0.05, seed=1, steps=100,
               precision="fp64", kmax=30, errtol=1e-8),
    "c2": dict(grid=(512, 128, 128), name="512x128x128", re_theta=600.0, ic="turbulent", intensity=0.10, seed=2,
               steps=500, precision="fp64", kmax=40, errtol=1e-10),
    "c3": dict(grid=(1024, 256, 192), name="1024x192x256", re_theta=1000.0, ic="turbulent", intensity=0.10,
               seed=3, steps=1000, precision="mixed", kmax=40, errtol=1e-10),
    "c4": dict(grid=(256, 64, 96), name="256x96x64", re_theta=400.0, ic="turbulent", intensity=0.10, seed=4,
               steps=200, precision="fp64", kmax=40, errtol=1e-10, nsnap=20),
    "c5a": dict(grid=(2048, 512, 256), name="2048x256x512", re_theta=1500.0, ic="turbulent", intensity=0.10,
                seed=5, steps=200, precision="mixed", kmax=40, errtol=1e-10),
    "c5b": dict(grid=(1024, 192, 160), name="1024x160x192", re_theta=1500.0, ic="turbulent", intensity=0.10,
                seed=5, steps=100, precision="mixed", kmax=40, errtol=1e-10, nsnap=10),
}


def stretch_delta(Lz=DOMAIN["Lz"], g=STRETCH):
    """the reference's tanh parameter delta for stretch factor g"""
    return 2.0 * g / Lz


def rec_plane(nX):
    return int(round(0.8 * (nX - 1))) + 1


def planes(nX):
    """the reference's mass-flux planes 5, 20, 40, 60, 85 (of nX = 90) scaled to nX"""
    return tuple(max(2, min(nX - 1, int(round(p / 90.0 * nX)))) for p in (5, 20, 40, 60, 85))


def theta_of(U, x3, Ue):
    """momentum thickness int (U/Ue)(1 - U/Ue) dz (trapezoidal)"""
    r = np.asarray(U) / Ue
    return float(np.trapezoid(r * (1.0 - r), np.asarray(x3)))


def turbulent_parameters(re_theta, kappa_cf=0.384, b_cf=4.127):
    """(Re, Ue) of a turbulent inflow with momentum-thickness Reynolds number re_theta"""
    Ue = math.log(re_theta) / kappa_cf + b_cf
    z = np.linspace(0.0, 1.0, 200001)

    def resid(Re):
        return theta_of(inflow.mean_profile(z, Re, Ue), z, Ue) * Re * Ue - re_theta

    lo, hi = 10.0, 1.0e5
    for _ in range(200):
        mid = math.sqrt(lo * hi)
        if resid(mid) > 0.0:
            hi = mid
        else:
            lo = mid
        if hi / lo < 1.0 + 1e-13:
            break
    return math.sqrt(lo * hi), Ue


def blasius_parameters(re_theta):
    """(Re, Ue) of a Blasius inflow with delta99 = 1 and momentum-thickness Reynolds number re_theta"""
    theta = BLASIUS_THETA_ETA / BLASIUS_ETA99
    re_delta = re_theta / theta
    Ue = math.sqrt(re_delta / (BLASIUS_FPP0 * BLASIUS_ETA99))
    return re_delta / Ue, Ue


def flow_parameters(case):
    c = CASES[case] if isinstance(case, str) else case
    if c["ic"] == "blasius":
        return blasius_parameters(c["re_theta"])
    return turbulent_parameters(c["re_theta"])


def blasius_fprime(eta):
    """f'(eta) of the Blasius solution (RK4 shooting on a fine grid, interpolated)"""
    h = 1e-3
    n = int(12.0 / h)
    y = np.array([0.0, 0.0, BLASIUS_FPP0])
    fp = np.empty(n + 1)
    fp[0] = 0.0

    def F(v):
        return np.array([v[1], v[2], -0.5 * v[0] * v[2]])

    for q in range(n):
        k1 = F(y)
        k2 = F(y + 0.5 * h * k1)
        k3 = F(y + 0.5 * h * k2)
        k4 = F(y + h * k3)
        y = y + (h / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4)
        fp[q + 1] = y[1]
    grid = np.arange(n + 1) * h
    return np.interp(np.asarray(eta, dtype=np.float64), grid, fp, right=1.0)


def blasius_profile(x3, Ue):
    return Ue * blasius_fprime(BLASIUS_ETA99 * np.asarray(x3, dtype=np.float64))


def blasius_field(mesh, Ue, noise=0.05, seed=1):
    """Ut (nX,nY,nZ,3): Blasius mean profile + uniform random U(-noise, noise) * Ue
    per component and node (Philox-seeded), no slip at the wall, periodic span ghosts"""
    x1, x2, x3 = (np.asarray(mesh[k], dtype=np.float64) for k in ("x1", "x2", "x3"))
    nX, nY, nZ = len(x1), len(x2), len(x3)
    rng = np.random.Generator(np.random.Philox(seed))
    buf = rng.uniform(-noise, noise, size=(3, nZ, nY, nX)) * Ue
    buf[0] += blasius_profile(x3, Ue)[:, None, None]
    buf[:, :, 0, :] = buf[:, :, nY - 2, :]
    buf[:, :, nY - 1, :] = buf[:, :, 1, :]
    buf[:, 0, :, :] = 0.0
    return buf.transpose(3, 2, 1, 0)


def cfl_dt(Ut, mesh, cfl=CFL):
    """dt with dt * max(|u|/dx + |v|/dy + |w|/dz_k) = cfl over the interior nodes"""
    x1, x2, x3 = (np.asarray(mesh[k], dtype=np.float64) for k in ("x1", "x2", "x3"))
    dx, dy = x1[1] - x1[0], x2[1] - x2[0]
    dz = 0.5 * (x3[2:] - x3[:-2])
    I = (slice(1, -1),) * 3
    a = (np.abs(Ut[I + (0,)]) / dx + np.abs(Ut[I + (1,)]) / dy
         + np.abs(Ut[I + (2,)]) / dz[None, None, :])
    return float(cfl / a.max())


def solver_params(case, dt, Re, Ue, grid=None, **over):
    """tbl.Solver keyword arguments of a case on its grid (or `grid`): domain,
    physics, recycling plane, mass-flux planes and solver settings"""
    c = CASES[case] if isinstance(case, str) else case
    nX, nY, nZ = grid or c["grid"]
    kw = dict(Lx=DOMAIN["Lx"], Ly=DOMAIN["Ly"], Lz=DOMAIN["Lz"], delta=stretch_delta(), dt=dt, Re=Re, Ue=Ue,
              i_rec=rec_plane(nX), planes=planes(nX), kmax=c["kmax"], errtol=c["errtol"])
    kw.update(over)
    return kw


def initial_state(case, mesh, Re, Ue):
    """(Ut, recycle state) of a case on the given mesh (dict x1, x2, x3)"""
    c = CASES[case] if isinstance(case, str) else case
    nX = len(mesh["x1"])
    i_rec = rec_plane(nX)
    if c["ic"] == "blasius":
        Ut = blasius_field(mesh, Ue, c.get("noise", 0.05), c["seed"])
        U = blasius_profile(mesh["x3"], Ue)
    else:
        Ut = inflow.initial_field(mesh, Re, Ue, seed=c["seed"], intensity=c.get("intensity", 0.10))
        U = inflow.mean_profile(mesh["x3"], Re, Ue)
    rs = inflow.recycle_state(mesh["x3"], nX, Re, Ue, i_rec, profile=U)
    return Ut, rs


def describe(case, mesh, Re, Ue, dt):
    """physical description of a case on its mesh (for records)"""
    c = CASES[case] if isinstance(case, str) else case
    x1, x3 = np.asarray(mesh["x1"]), np.asarray(mesh["x3"])
    dx = x1[1] - x1[0]
    return dict(re_theta_in=c["re_theta"], Re_tau=round(Re, 3), Ue=round(Ue, 4), dt=dt,
                cfl=CFL, cfl_x_free_stream=round(Ue * dt / dx, 4),
                domain="%gx%gx%g delta99 (stream x wall x span)" % (DOMAIN["Lx"], DOMAIN["Lz"], DOMAIN["Ly"]),
                stretch=STRETCH, first_cell_yplus=round(float((x3[1] - x3[0]) * Re), 3), ic=c["ic"])
