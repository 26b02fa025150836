"""Seeded initial fields and recycling state for the TBL step.

The reference reads its inflow mean profile from ``Input/initProf.dat`` and its
restart state from ``Input/3D.dat`` / ``Input/statData.dat``
(dnssetbl/dnssetbl.f90:106-150); none of these files ship with it.  This
module builds the same objects from a formula and a seed instead:

* ``mean_profile``   -- Reichardt inner law + Coles wake in wall units
  (u_tau = 1; lengths in inflow boundary-layer thicknesses, Re = u_tau*delta/nu),
  reaching ``Ue`` at z = delta99 and constant above it.
* ``initial_field``  -- mean profile + divergence-free random Fourier modes
  (curl of a vector potential; streamwise/spanwise wavenumber indices <= kmax,
  damped to zero with zero slope at the wall; RMS = intensity * Ue in the most
  intense wall-normal plane), Philox-seeded so that the field does not depend on any
  decomposition.
* ``recycle_state``  -- the statistics the recycling inflow reads
  (usum2/vsum2/wsum2 at the recycling plane, Wt(1,:), new(k), lambda) set as if
  a previous averaging window had produced the mean profile, i.e. the state a
  restart with nstep > freqTimAvg starts from (dnssetbl.f90:118-150).
"""
import math

import numpy as np

KAPPA = 0.41


def reichardt(zp):
    zp = np.asarray(zp, dtype=np.float64)
    return (np.log1p(KAPPA * zp) / KAPPA
            + 7.8 * (1.0 - np.exp(-zp / 11.0) - (zp / 11.0) * np.exp(-zp / 3.0)))


def mean_profile(x3, Re, Ue, delta99=1.0):
    """U(z) in units of u_tau; wake strength chosen so that U(delta99) = Ue."""
    x3 = np.asarray(x3, dtype=np.float64)
    u_edge = float(reichardt(delta99 * Re))
    wake = max(Ue - u_edge, 0.0)
    z = np.minimum(x3, delta99)
    u = reichardt(z * Re) + wake * np.sin(0.5 * math.pi * z / delta99) ** 2
    return np.where(x3 >= delta99, Ue, np.minimum(u, Ue))


def _modes(rng, nmodes, kmax):
    kx = rng.integers(1, kmax + 1, size=nmodes)
    ky = rng.integers(1, kmax + 1, size=nmodes)
    amp = rng.standard_normal((nmodes, 3))
    ph = rng.uniform(0.0, 2.0 * math.pi, size=(nmodes, 3, 2))
    return kx, ky, amp, ph


def initial_field(mesh, Re, Ue, seed=2, intensity=0.10, kmax=16, nmodes=256,
                  delta99=1.0, zdamp=0.5):
    """Ut (nX,nY,nZ,3) float64 Fortran-ordered: mean + div-free perturbation.

    ``mesh`` holds x1 (nX), x2 (nY), x3 (nZ).  The span period is the
    interior period (nY-2)*dy of the reference's periodic tables; the
    streamwise modes use the domain length.  The perturbation is the curl of
    A = sum_m a_m cos(kx x + p) cos(ky y + q) f(z), f = (z/zd)^2 exp(-(z/zd)^2),
    evaluated analytically so that it is exactly solenoidal in the continuum;
    kx, ky are mode indices 1..kmax.  ``intensity`` is the RMS of the three
    perturbation components over the interior nodes of the most intense
    wall-normal plane, in units of Ue.

    Every mode shares f(z), so each component is f(z) P(x,y) + f'(z) R(x,y)
    with plane sums P, R over the modes: O(nmodes nX nY + N) work.
    """
    x1 = np.asarray(mesh["x1"], dtype=np.float64)
    x2 = np.asarray(mesh["x2"], dtype=np.float64)
    x3 = np.asarray(mesh["x3"], dtype=np.float64)
    nX, nY, nZ = len(x1), len(x2), len(x3)
    Lx = x1[-1] - x1[0]
    dy = x2[1] - x2[0]
    Ly = (nY - 2) * dy
    rng = np.random.Generator(np.random.Philox(seed))
    kx, ky, amp, ph = _modes(rng, nmodes, kmax)
    s = x3 / zdamp
    f = s * s * np.exp(-s * s)
    fp = (2.0 * s * np.exp(-s * s) * (1.0 - s * s)) / zdamp
    ax = 2.0 * math.pi * kx / Lx                                 # (nmodes,)
    ay = 2.0 * math.pi * ky / Ly
    # per mode and potential component q: X_q(x), X_q'(x), Y_q(y), Y_q'(y)
    X = np.cos(ax[:, None, None] * x1[None, None, :] + ph[:, :, 0, None])       # (m, 3, nX)
    Xd = -ax[:, None, None] * np.sin(ax[:, None, None] * x1[None, None, :] + ph[:, :, 0, None])
    yy = (x2 - x2[1])[None, None, :]
    Y = np.cos(ay[:, None, None] * yy + ph[:, :, 1, None])                      # (m, 3, nY)
    Yd = -ay[:, None, None] * np.sin(ay[:, None, None] * yy + ph[:, :, 1, None])

    def plane(q, xs, ys):
        # sum_m a_mq ys_mq(y) xs_mq(x) -> (nY, nX)
        return np.einsum("m,my,mx->yx", amp[:, q], ys[:, q, :], xs[:, q, :], optimize=True)

    # A_q = a_q X_q(x) Y_q(y) f(z); u' = dA3/dy - dA2/dz ; v' = dA1/dz - dA3/dx ; w' = dA2/dx - dA1/dy
    terms = [(plane(2, X, Yd), -plane(1, X, Y)),                 # u': f * P + f' * R
             (-plane(2, Xd, Y), plane(0, X, Y)),                 # v'
             (plane(1, Xd, Y) - plane(0, X, Yd), None)]          # w'
    # C-ordered (3, nZ, nY, nX) buffer == Fortran-ordered (nX, nY, nZ, 3)
    buf = np.empty((3, nZ, nY, nX))
    for c, (P, R) in enumerate(terms):
        np.multiply(f[:, None, None], P[None, :, :], out=buf[c])
        if R is not None:
            buf[c] += fp[:, None, None] * R[None, :, :]
    Ut = buf.transpose(3, 2, 1, 0)
    assert Ut.flags.f_contiguous
    # RMS per wall-normal plane over the interior (x, y) nodes and the three
    # components; the most intense plane is scaled to intensity * Ue
    sub = buf[:, 1:-1, 1:-1, 1:-1]
    rms = math.sqrt(float(np.max(np.mean(sub * sub, axis=(0, 2, 3))))) if sub.size else 0.0
    if rms > 0:
        buf *= intensity * Ue / rms
    buf[0] += mean_profile(x3, Re, Ue, delta99)[:, None, None]
    # periodic span ghosts and no-slip wall, as the reference's BCs leave them
    buf[:, :, 0, :] = buf[:, :, nY - 2, :]
    buf[:, :, nY - 1, :] = buf[:, :, 1, :]
    buf[:, 0, :, :] = 0.0
    return Ut


def lund_weight(eta):
    """Lund et al. (1998) inner/outer blending weight as written at dnssetbl.f90:1082."""
    eta = np.asarray(eta, dtype=np.float64)
    return 0.5 * (1 + (np.tanh((4.0 * (eta - 0.2)) / ((1 - 2.0 * 0.2) * eta + 0.2)) / np.tanh(4.0)))


def recycle_state(x3, nX, Re, Ue, i_rec, delta99=1.0, profile=None):
    """Statistics a restart would read: see module docstring.

    Returns dict(avg (nX,nZ,9), Wt (nX,nZ), newk (nZ,) int32 1-based, lamda,
    lamda03, Ume (nZ,)).  With lambda = lambda03 = 1 the map new(k) is the
    identity on k = 1..nZ-1 (the reference's scan, dnssetbl.f90:1096-1129).
    ``profile`` (nZ,) replaces the turbulent mean profile (e.g. a Blasius one).
    """
    x3 = np.asarray(x3, dtype=np.float64)
    nZ = len(x3)
    U = mean_profile(x3, Re, Ue, delta99) if profile is None else np.asarray(profile, dtype=np.float64)
    avg = np.zeros((nX, nZ, 9), order="F")
    avg[:, :, 0] = U[None, :]
    avg[:, :, 3] = (U * U)[None, :]
    # delta99 of the profile as the reference finds it (first node U >= 0.99 Ue)
    k99 = int(np.argmax(U >= 0.99 * Ue))
    bl = x3[k99]
    Wt = np.asfortranarray(np.broadcast_to(lund_weight(x3 / bl)[None, :], (nX, nZ)).copy())
    newk = np.full(nZ, nZ - 1, dtype=np.int32)
    newk[: nZ - 1] = np.arange(1, nZ, dtype=np.int32)
    return dict(avg=avg, Wt=Wt, newk=newk, lamda=1.0, lamda03=1.0, Ume=U.copy(),
                bltkns=np.full(nX, bl))
