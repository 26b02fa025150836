"""ctypes wrapper of the CPU oracle (oracle/tbl_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The shipped package
(turbulentbl_amd) never imports it.

All field arrays are float64, Fortran-ordered, with the reference's shapes:
Ut (nX,nY,nZ,3), Press/PCorr/pRhs (nX,nY,nZ), statistics (nX,nZ).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborc.so")

PC_SSOR_BLOCK, PC_SSOR_RB, PC_NONE, PC_JACOBI = 0, 1, 2, 3
GS_LEX_BLOCK, GS_RB, GS_ZLINE_RB, GS_ZEBRA = 0, 1, 2, 3

# the reference's SSOR factor is the REAL(4) literal 1.835 (dnssetbl.f90:34)
OMEGA_REF = float(np.float32(1.835))
REF_PLANES = (5, 20, 40, 60, 85)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    P = C.c_void_p
    d, i = C.c_double, C.c_int
    lib.orc_grid_create.restype = P
    lib.orc_grid_create.argtypes = [i, i, i, i, d, d, d, d, d, d, d, d, i, i, P]
    lib.orc_grid_set_scheme.restype = None
    lib.orc_grid_set_scheme.argtypes = [P, i, P]
    lib.orc_grid_destroy.argtypes = [P]
    lib.orc_grid_mesh.argtypes = [P] * 12
    lib.orc_grid_scalars.argtypes = [P, P]
    lib.orc_recycle_inflow.argtypes = [P, i, d, P, P, P, P, P, P, P, P]
    lib.orc_momentum_rhs.argtypes = [P, P, P, P, P]
    lib.orc_diffusion_solve.argtypes = [P, P, P, i, d, i, P, P]
    lib.orc_helm_residual.argtypes = [P, P, P, i]
    lib.orc_helm_residual.restype = d
    lib.orc_helm_sweep.argtypes = [P, P, P, i, i]
    lib.orc_ppe_rhs.argtypes = [P, P, P]
    lib.orc_poisson_apply.argtypes = [P, P, P]
    lib.orc_precond_apply.argtypes = [P, i, P, P]
    lib.orc_ppe_gmres.argtypes = [P, P, P, i, d, i, i, P]
    lib.orc_project.argtypes = [P, P, P, P]
    lib.orc_apply_bcs.argtypes = [P, P, P, P, P]
    lib.orc_div_check.argtypes = [P, P, P]
    lib.orc_div_check.restype = d
    lib.orc_mass_flux.argtypes = [P, P, P]
    lib.orc_stats_accumulate.argtypes = [P, P, P]
    lib.orc_rescale.argtypes = [P, P, i, P, P, P, P, P, P, P, P, P, P]
    lib.orc_ke_guard.argtypes = [P, P, P, P, P, P, P]
    lib.orc_ke_guard.restype = d
    lib.orc_wall_shear.argtypes = [P, P, P]
    lib.orc_zline_chunks.argtypes = [i]
    lib.orc_zline_begin.argtypes = [i, i, i]
    lib.orc_threads.restype = i
    lib.orc_threads.argtypes = []
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def zline_chunks(nz):
    """chunks of the z-line partition for nz = nZ-2 interior planes"""
    return int(lib().orc_zline_chunks(nz))


def zline_begin(ch, C, nz):
    """first (1-based) plane of chunk ch"""
    return int(lib().orc_zline_begin(ch, C, nz))


def _p(a):
    """Pointer to a contiguous Fortran-ordered array (or None)."""
    if a is None:
        return None
    assert a.flags.f_contiguous, "oracle arrays must be Fortran-ordered"
    return a.ctypes.data


def f64(a):
    return np.asfortranarray(a, dtype=np.float64)


class Grid:
    """The reference's mesh, metrics and coefficient set (dnssetbl.f90:183-247)."""

    def __init__(self, nX, nY, nZ, P=1, Lx=12.0, Ly=4.0, Lz=3.4, delta=1.95,
                 dt=1.0e-4, Re=150.92, Ue=19.798, omega=OMEGA_REF, i_rec=68,
                 freq_tim_avg=500, planes=REF_PLANES, scheme=0, forc=(0.0, 0.0, 0.0)):
        self.nX, self.nY, self.nZ, self.P = nX, nY, nZ, P
        self.Lx, self.Ly, self.Lz, self.delta = Lx, Ly, Lz, delta
        self.scheme, self.forc = int(scheme), tuple(float(f) for f in forc)
        self.dt, self.Re, self.Ue, self.omega = dt, Re, Ue, omega
        self.i_rec, self.freq_tim_avg = i_rec, freq_tim_avg
        self.planes = tuple(int(p) for p in planes)
        pl = (C.c_int * 5)(*self.planes)
        self.h = lib().orc_grid_create(nX, nY, nZ, P, Lx, Ly, Lz, delta, dt, Re, Ue,
                                       omega, i_rec, freq_tim_avg, pl)
        if not self.h:
            raise ValueError("invalid oracle grid")
        fc = (C.c_double * 3)(*self.forc)
        lib().orc_grid_set_scheme(self.h, self.scheme, C.cast(fc, C.c_void_p))
        m = {k: np.zeros(nZ) for k in ("x3", "Xi", "Jac", "Jac2", "JacInv", "Ab", "At", "Ap", "rAp")}
        m["x1"], m["x2"] = np.zeros(nX), np.zeros(nY)
        lib().orc_grid_mesh(self.h, *[m[k].ctypes.data for k in
                                      ("x1", "x2", "x3", "Xi", "Jac", "Jac2", "JacInv", "Ab", "At", "Ap", "rAp")])
        self.mesh = m
        s = np.zeros(9)
        lib().orc_grid_scalars(self.h, s.ctypes.data)
        (self.dx, self.dy, self.dXi, self.Aw, self.Ae, self.As, self.An,
         self.df1, self.df2) = s.tolist()

    def __del__(self):
        try:
            if self.h:
                lib().orc_grid_destroy(self.h)
        except Exception:
            pass

    # shapes
    @property
    def N(self):
        return self.nX * self.nY * self.nZ

    def zeros(self, *extra):
        return np.zeros((self.nX, self.nY, self.nZ) + extra, order="F")

    def zeros2(self, *extra):
        return np.zeros((self.nX, self.nZ) + extra, order="F")

    # ---- blocks ----
    def recycle_inflow(self, Ut, nstep, lamda, avg, Wt, newk, Ume, Ufluc_rec):
        newk = np.ascontiguousarray(newk, dtype=np.int32)
        lib().orc_recycle_inflow(self.h, nstep, lamda, _p(avg[..., 0]), _p(avg[..., 1]), _p(avg[..., 2]),
                                 _p(Wt), newk.ctypes.data, f64(Ume).ctypes.data, _p(Ufluc_rec), _p(Ut))

    def momentum_rhs(self, Ut, Press):
        conv, Q = self.zeros(3), self.zeros(3)
        lib().orc_momentum_rhs(self.h, _p(Ut), _p(Press), _p(conv), _p(Q))
        return conv, Q

    def diffusion_solve(self, Q, Ut, order=GS_LEX_BLOCK, errorstop=1.0e-6, mmax=1000):
        sw = np.zeros(3, dtype=np.int32)
        res = np.zeros(3)
        lib().orc_diffusion_solve(self.h, _p(Q), _p(Ut), order, errorstop, mmax,
                                  sw.ctypes.data, res.ctypes.data)
        return sw, res

    def helm_residual(self, Q, Ut, p):
        return lib().orc_helm_residual(self.h, _p(Q), _p(Ut), p)

    def helm_sweep(self, Q, Ut, p, order=GS_LEX_BLOCK):
        lib().orc_helm_sweep(self.h, _p(Q), _p(Ut), p, order)

    def ppe_rhs(self, Ut):
        out = self.zeros()
        lib().orc_ppe_rhs(self.h, _p(Ut), _p(out))
        return out

    def poisson_apply(self, x):
        out = self.zeros()
        lib().orc_poisson_apply(self.h, _p(f64(x)), _p(out))
        return out

    def precond_apply(self, r, pc=PC_SSOR_BLOCK):
        out = self.zeros()
        lib().orc_precond_apply(self.h, pc, _p(f64(r)), _p(out))
        return out

    def ppe_gmres(self, pRhs, PCorr, pc=PC_SSOR_BLOCK, errtol=1.0e-5, kmax=50, mmax=1000):
        st = np.zeros(6)
        lib().orc_ppe_gmres(self.h, _p(pRhs), _p(PCorr), pc, errtol, kmax, mmax, st.ctypes.data)
        return dict(cycles=int(st[0]), iters=int(st[1]), last_kit=int(st[2]), sumav=st[3],
                    sumnor=st[4], true_res=st[5])

    def project(self, PCorr, Press, Ut):
        lib().orc_project(self.h, _p(PCorr), _p(Press), _p(Ut))

    def apply_bcs(self, Ut, Utold, Press, wall_w=None):
        lib().orc_apply_bcs(self.h, _p(Ut), _p(Utold), _p(Press),
                            _p(f64(wall_w)) if wall_w is not None else None)

    def div_check(self, Ut):
        s = C.c_double(0.0)
        m = lib().orc_div_check(self.h, _p(Ut), C.addressof(s))
        return m, s.value

    def mass_flux(self, Ut):
        out = np.zeros(5)
        lib().orc_mass_flux(self.h, _p(Ut), out.ctypes.data)
        return out

    def stats_accumulate(self, Ut, sums):
        lib().orc_stats_accumulate(self.h, _p(Ut), _p(sums))

    def rescale(self, sums, nsamples, Ut, bltkns=None):
        avg = self.zeros2(9)
        Wt = self.zeros2()
        newk = np.zeros(self.nZ, dtype=np.int32)
        bl = np.zeros(self.nX) if bltkns is None else np.array(bltkns, dtype=np.float64)
        ds, th = np.zeros(self.nX), np.zeros(self.nX)
        uf = np.zeros((self.nY, self.nZ, 3), order="F")
        lam, lam03 = C.c_double(0), C.c_double(0)
        lib().orc_rescale(self.h, _p(sums), nsamples, _p(Ut), _p(avg), C.addressof(lam),
                          C.addressof(lam03), _p(Wt), newk.ctypes.data, bl.ctypes.data,
                          ds.ctypes.data, th.ctypes.data, _p(uf))
        return dict(avg=avg, lamda=lam.value, lamda03=lam03.value, Wt=Wt, newk=newk,
                    bltkns=bl, deltastar=ds, theta=th, Ufluc_rec=uf)

    def ke_guard(self, Ut, avg, Ufluc_rec=None):
        gd = C.c_int(0)
        ke = lib().orc_ke_guard(self.h, _p(Ut), _p(avg[..., 0]), _p(avg[..., 1]), _p(avg[..., 2]),
                                C.addressof(gd), _p(Ufluc_rec))
        return ke, gd.value

    def wall_shear(self, Ut):
        tau = np.zeros(self.nX)
        lib().orc_wall_shear(self.h, _p(Ut), tau.ctypes.data)
        return tau


class State:
    """Mutable solver state of the oracle's whole-step driver (mirrors orc_state)."""

    FIELDS = ("Ut", "Utold", "Press", "PCorr", "pRhs", "conv", "Q")

    def __init__(self, g: Grid):
        self.g = g
        self.Ut, self.Utold, self.conv, self.Q = g.zeros(3), g.zeros(3), g.zeros(3), g.zeros(3)
        self.Press, self.PCorr, self.pRhs = g.zeros(), g.zeros(), g.zeros()
        self.sums, self.avg = g.zeros2(9), g.zeros2(9)
        self.Wt = g.zeros2()
        self.newk = np.full(g.nZ, g.nZ - 1, dtype=np.int32)
        self.Ume = np.zeros(g.nZ)
        self.Ufluc_rec = np.zeros((g.nY, g.nZ, 3), order="F")
        self.bltkns, self.deltastar, self.theta = np.zeros(g.nX), np.zeros(g.nX), np.zeros(g.nX)
        self.wall_w = None
        # wall-shear window (dnstbl/dnstbl.f90:1666-1682): 0 = off
        self.tau_window = 0
        self.tausum, self.taumean, self.cf = np.zeros(g.nX), np.zeros(g.nX), np.zeros(g.nX)
        self.nwin = 0
        self.lamda = 1.0
        self.lamda03 = 1.0
        self.time = 0.0
        self.nstep = 0
        self.cnt = 0

    def step(self, gs_order=GS_LEX_BLOCK, errorstop=1.0e-6, gs_mmax=1000,
             pc=PC_SSOR_BLOCK, errtol=1.0e-5, kmax=50, mmax=1000):
        """One reference time step (dnssetbl.f90:253-1262); returns diagnostics."""
        g = self.g
        self.time += g.dt
        self.nstep += 1
        self.cnt += 1
        self.Utold[...] = self.Ut
        g.recycle_inflow(self.Ut, self.nstep, self.lamda, self.avg, self.Wt, self.newk,
                         self.Ume, self.Ufluc_rec)
        self.conv, self.Q = g.momentum_rhs(self.Ut, self.Press)
        sweeps, gsres = g.diffusion_solve(self.Q, self.Ut, gs_order, errorstop, gs_mmax)
        self.pRhs = g.ppe_rhs(self.Ut)
        gm = g.ppe_gmres(self.pRhs, self.PCorr, pc, errtol, kmax, mmax)
        g.project(self.PCorr, self.Press, self.Ut)
        g.apply_bcs(self.Ut, self.Utold, self.Press, self.wall_w)
        div, _ = g.div_check(self.Ut)
        mass = g.mass_flux(self.Ut)
        cf_ready = False
        if self.tau_window > 0:
            # every step's wall shear into the window sum; at the window's last
            # step the mean sum/window per station as tau/(nY-2) and
            # C_f = 2 (sum/window) / (Ue^2 (nY-2)) (dnstbl.f90:1666-1682)
            tau = g.wall_shear(self.Ut)
            self.tausum[1:-1] = tau[1:-1] + self.tausum[1:-1]
            if self.nstep % self.tau_window == 0:
                sw = self.tausum / float(self.tau_window)
                self.taumean = sw / float(g.nY - 2)
                self.cf = 2.0 * (sw / (g.Ue * g.Ue * float(g.nY - 2)))
                self.tausum = np.zeros(g.nX)
                self.nwin += 1
                cf_ready = True
        g.stats_accumulate(self.Ut, self.sums)
        rescaled = False
        if self.nstep % g.freq_tim_avg == 0:
            r = g.rescale(self.sums, g.freq_tim_avg, self.Ut, self.bltkns)
            self.avg = r["avg"]
            self.lamda, self.lamda03 = r["lamda"], r["lamda03"]
            self.Wt, self.newk, self.bltkns = r["Wt"], r["newk"], r["bltkns"]
            self.deltastar, self.theta = r["deltastar"], r["theta"]
            uf = r["Ufluc_rec"]
            # the recycling-plane fluctuations are refreshed on k=2..nZ-1
            self.Ufluc_rec[:, 1:-1, :] = uf[:, 1:-1, :]
            self.sums[...] = 0.0
            rescaled = True
        ke, guard = 0.0, 0
        if self.cnt > g.freq_tim_avg:
            ke, guard = g.ke_guard(self.Ut, self.avg, self.Ufluc_rec)
        return dict(div=div, mass=mass, ke=ke, guard=guard, gs_sweeps=sweeps, gs_res=gsres,
                    gmres=gm, rescaled=rescaled, cf_ready=cf_ready)
