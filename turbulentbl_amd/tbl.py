"""Python host binding of libtbl_hip.so (the C ABI in include/tbl.h).

The reference's time step (dnssetbl/dnssetbl.f90:253-1262) is a sequence of
inline blocks over global arrays; :class:`Solver` exposes one method per block
with the reference's array shapes (Fortran-ordered float64, Ut (nX,nY,nZ,3),
Press/PCorr (nX,nY,nZ), statistics (nX,nZ)), plus ``step`` for whole steps.
All compute runs in the HIP library on the GPU; there is no CPU fallback:
loading fails loudly if the library is missing or a GPU is absent.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtbl_hip.so")

TBL_FP64, TBL_MIXED = 0, 1
TBL_PC_SSOR_RB, TBL_PC_NONE, TBL_PC_JACOBI, TBL_PC_FAST, TBL_PC_SSOR_BLOCK = 1, 2, 3, 4, 5
TBL_HELM_POINT_RB, TBL_HELM_ZLINE_RB, TBL_HELM_DIRECT, TBL_HELM_ZEBRA, TBL_HELM_LEX_SLAB = 0, 1, 2, 3, 4
TBL_SCHEME_SETBL, TBL_SCHEME_DNSTBL = 0, 1
TBL_TOL_REF, TBL_TOL_TRUE = 0, 1
F_UT, F_UTOLD, F_PRESS, F_PCORR, F_PRHS, F_Q, F_WORK = range(7)
KCLASSES = ("rhs", "helm", "helm_res", "div", "poisson", "precond", "dots", "update", "xupd",
            "resid", "project", "bcs", "diag", "recycle", "rescale")
ERRORS = {-1: "TBL_EARG", -2: "TBL_EBLOWUP", -3: "TBL_ENAN", -4: "TBL_ENOCONV", -5: "TBL_ESTATE"}


class TblError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "HIP error"), code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [("nX", C.c_int), ("nY", C.c_int), ("nZ", C.c_int),
                ("Lx", C.c_double), ("Ly", C.c_double), ("Lz", C.c_double), ("delta", C.c_double),
                ("dt", C.c_double), ("Re", C.c_double), ("Ue", C.c_double), ("omega", C.c_double),
                ("i_rec", C.c_int), ("freq_tim_avg", C.c_int), ("planes", C.c_int * 5),
                ("precision", C.c_int), ("errorstop", C.c_double), ("gs_mmax", C.c_int), ("helm_mode", C.c_int),
                ("tol_mode", C.c_int), ("errtol", C.c_double), ("kmax", C.c_int), ("mmax", C.c_int),
                ("precond", C.c_int), ("inner_tol", C.c_double), ("fail_on_noconv", C.c_int),
                ("wall_shear_window", C.c_int), ("pc_extrap", C.c_int), ("scheme", C.c_int),
                ("forc", C.c_double * 3), ("slabs", C.c_int)]


class State(C.Structure):
    _fields_ = [("sums", C.c_void_p), ("avg", C.c_void_p), ("Wt", C.c_void_p), ("newk", C.c_void_p),
                ("Ume", C.c_void_p), ("Ufluc_rec", C.c_void_p), ("bltkns", C.c_void_p),
                ("deltastar", C.c_void_p), ("theta", C.c_void_p), ("wall_w", C.c_void_p),
                ("lamda", C.c_double), ("lamda03", C.c_double), ("time", C.c_double),
                ("nstep", C.c_int), ("cnt", C.c_int)]


class Diag(C.Structure):
    _fields_ = [("div", C.c_double), ("div_signed", C.c_double), ("mass", C.c_double * 5),
                ("ke", C.c_double), ("guard", C.c_int), ("gs_sweeps", C.c_int * 3),
                ("gm_cycles", C.c_int), ("gm_iters", C.c_int), ("gm_true_res", C.c_double),
                ("gm_sumav", C.c_double), ("rescaled", C.c_int), ("nstep", C.c_int), ("cf_ready", C.c_int)]

    def as_dict(self):
        return dict(div=self.div, div_signed=self.div_signed, mass=list(self.mass), ke=self.ke,
                    guard=self.guard, gs_sweeps=list(self.gs_sweeps), gm_cycles=self.gm_cycles,
                    gm_iters=self.gm_iters, gm_true_res=self.gm_true_res, gm_sumav=self.gm_sumav,
                    rescaled=self.rescaled, nstep=self.nstep, cf_ready=self.cf_ready)


class AdjointOpts(C.Structure):
    _fields_ = [("nsteps", C.c_int), ("nsnap", C.c_int), ("i0", C.c_int), ("i1", C.c_int),
                ("forward_only", C.c_int), ("branch_mode", C.c_int)]


EXPORTS = ("tbl_default_params", "tbl_create", "tbl_destroy", "tbl_last_error", "tbl_sync",
           "tbl_device_bytes", "tbl_upload", "tbl_download", "tbl_checkpoint_save", "tbl_checkpoint_load", "tbl_field_set", "tbl_field_get",
           "tbl_set_state", "tbl_get_state", "tbl_get_mesh", "tbl_begin_step", "tbl_recycle_inflow",
           "tbl_momentum_rhs", "tbl_diffusion_solve", "tbl_ppe_rhs", "tbl_ppe_solve", "tbl_project",
           "tbl_apply_bcs", "tbl_div_check", "tbl_mass_flux", "tbl_stats_accumulate", "tbl_rescale",
           "tbl_ke_guard", "tbl_wall_shear", "tbl_wall_shear_mean", "tbl_div_check_noalloc", "tbl_step", "tbl_apply_operator", "tbl_timer_enable",
           "tbl_timer_read", "tbl_adjoint_window")

_lib = None


def load(path=LIB_PATH):
    """dlopen the HIP library (no device call).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError("libtbl_hip.so is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    lib.tbl_default_params.argtypes = [C.POINTER(Params)]
    lib.tbl_default_params.restype = None
    lib.tbl_create.argtypes = [C.POINTER(Params), C.POINTER(P)]
    lib.tbl_destroy.argtypes = [P]
    lib.tbl_last_error.argtypes = [P, C.c_char_p, I]
    lib.tbl_sync.argtypes = [P]
    lib.tbl_device_bytes.argtypes = [P]
    lib.tbl_device_bytes.restype = D
    lib.tbl_upload.argtypes = [P, P, P, P]
    lib.tbl_download.argtypes = [P, P, P, P]
    lib.tbl_field_set.argtypes = [P, I, P]
    lib.tbl_field_get.argtypes = [P, I, P]
    lib.tbl_set_state.argtypes = [P, C.POINTER(State)]
    lib.tbl_get_state.argtypes = [P, C.POINTER(State)]
    lib.tbl_get_mesh.argtypes = [P] * 9
    for n in ("tbl_begin_step", "tbl_recycle_inflow", "tbl_momentum_rhs", "tbl_ppe_rhs", "tbl_project",
              "tbl_apply_bcs", "tbl_stats_accumulate", "tbl_rescale"):
        getattr(lib, n).argtypes = [P]
    lib.tbl_diffusion_solve.argtypes = [P, D, I, P, P]
    lib.tbl_ppe_solve.argtypes = [P, D, I, I, P, P, P]
    lib.tbl_div_check.argtypes = [P, P, P]
    lib.tbl_mass_flux.argtypes = [P, P]
    lib.tbl_ke_guard.argtypes = [P, P, P]
    lib.tbl_wall_shear.argtypes = [P, P]
    lib.tbl_checkpoint_save.argtypes = [P, C.c_char_p]
    lib.tbl_checkpoint_load.argtypes = [P, C.c_char_p]
    lib.tbl_wall_shear_mean.argtypes = [P, P, P, P]
    lib.tbl_div_check_noalloc.argtypes = [P, P, P]
    lib.tbl_step.argtypes = [P, I, P]
    lib.tbl_apply_operator.argtypes = [P, I, P, P]
    lib.tbl_timer_enable.argtypes = [P, C.c_uint]
    lib.tbl_timer_read.argtypes = [P, P, P, P]
    lib.tbl_adjoint_window.argtypes = [P, C.POINTER(AdjointOpts), P, P, P, P, P]
    for n in EXPORTS:
        if n not in ("tbl_default_params", "tbl_device_bytes"):
            getattr(lib, n).restype = I
    _lib = lib
    return lib


def default_params(**kw):
    p = Params()
    load().tbl_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "planes":
            p.planes = (C.c_int * 5)(*[int(x) for x in v])
        elif k == "forc":
            p.forc = (C.c_double * 3)(*[float(x) for x in v])
        elif k == "precision" and isinstance(v, str):
            p.precision = {"fp64": TBL_FP64, "mixed": TBL_MIXED}[v]
        else:
            setattr(p, k, v)
    return p


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def _ptr(a):
    if a is None:
        return None
    assert a.flags.f_contiguous and a.dtype in (np.float64, np.int32)
    return a.ctypes.data


class Solver:
    """One GPU context holding the whole domain (one HIP stream)."""

    def __init__(self, nX, nY, nZ, **kw):
        self.lib = load()
        self.p = default_params(nX=nX, nY=nY, nZ=nZ, **kw)
        h = C.c_void_p()
        rc = self.lib.tbl_create(C.byref(self.p), C.byref(h))
        if rc != 0:
            buf = C.create_string_buffer(512)
            self.lib.tbl_last_error(None, buf, 512)
            raise TblError(rc, buf.value.decode(errors="replace") or "tbl_create failed")
        self.h = h
        self.nX, self.nY, self.nZ = nX, nY, nZ
        self.N = nX * nY * nZ

    # -- plumbing
    def _chk(self, rc):
        if rc != 0:
            buf = C.create_string_buffer(512)
            self.lib.tbl_last_error(self.h, buf, 512)
            raise TblError(rc, buf.value.decode(errors="replace"))

    def close(self):
        if getattr(self, "h", None):
            self.lib.tbl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def zeros(self, *extra):
        return np.zeros((self.nX, self.nY, self.nZ) + extra, order="F")

    @property
    def device_bytes(self):
        return self.lib.tbl_device_bytes(self.h)

    def sync(self):
        self._chk(self.lib.tbl_sync(self.h))

    # -- data
    def upload(self, Ut=None, Press=None, PCorr=None):
        a = [None if x is None else _f64(x) for x in (Ut, Press, PCorr)]
        self._chk(self.lib.tbl_upload(self.h, *[_ptr(x) for x in a]))

    def download(self):
        Ut, P, PC = self.zeros(3), self.zeros(), self.zeros()
        self._chk(self.lib.tbl_download(self.h, _ptr(Ut), _ptr(P), _ptr(PC)))
        return Ut, P, PC

    def checkpoint_save(self, path):
        self._chk(self.lib.tbl_checkpoint_save(self.h, os.fsencode(path)))

    def checkpoint_load(self, path):
        self._chk(self.lib.tbl_checkpoint_load(self.h, os.fsencode(path)))

    def set_field(self, f, a):
        self._chk(self.lib.tbl_field_set(self.h, f, _ptr(_f64(a))))

    def get_field(self, f):
        out = self.zeros(3) if f in (F_UT, F_UTOLD, F_Q) else self.zeros()
        self._chk(self.lib.tbl_field_get(self.h, f, _ptr(out)))
        return out

    def set_state(self, sums=None, avg=None, Wt=None, newk=None, Ume=None, Ufluc_rec=None, bltkns=None,
                  deltastar=None, theta=None, wall_w=None, lamda=1.0, lamda03=1.0, time=0.0, nstep=0, cnt=0):
        keep = []

        def ref(a, dt=np.float64):
            if a is None:
                return None
            a = np.asfortranarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        s = State(ref(sums), ref(avg), ref(Wt), ref(newk, np.int32), ref(Ume), ref(Ufluc_rec), ref(bltkns),
                  ref(deltastar), ref(theta), ref(wall_w), lamda, lamda03, time, nstep, cnt)
        self._chk(self.lib.tbl_set_state(self.h, C.byref(s)))

    def get_state(self):
        nX, nY, nZ = self.nX, self.nY, self.nZ
        out = dict(sums=np.zeros((nX, nZ, 9), order="F"), avg=np.zeros((nX, nZ, 9), order="F"),
                   Wt=np.zeros((nX, nZ), order="F"), newk=np.zeros(nZ, dtype=np.int32),
                   Ume=np.zeros(nZ), Ufluc_rec=np.zeros((nY, nZ, 3), order="F"), bltkns=np.zeros(nX),
                   deltastar=np.zeros(nX), theta=np.zeros(nX), wall_w=np.zeros((nX, nY), order="F"))
        s = State(*[out[k].ctypes.data for k in ("sums", "avg", "Wt", "newk", "Ume", "Ufluc_rec", "bltkns",
                                                   "deltastar", "theta", "wall_w")], 0.0, 0.0, 0.0, 0, 0)
        self._chk(self.lib.tbl_get_state(self.h, C.byref(s)))
        out.update(lamda=s.lamda, lamda03=s.lamda03, time=s.time, nstep=s.nstep, cnt=s.cnt)
        return out

    def mesh(self):
        nX, nY, nZ = self.nX, self.nY, self.nZ
        m = dict(x1=np.zeros(nX), x2=np.zeros(nY), x3=np.zeros(nZ), Jac=np.zeros(nZ), Jac2=np.zeros(nZ),
                 Ab=np.zeros(nZ), At=np.zeros(nZ), Ap=np.zeros(nZ))
        self._chk(self.lib.tbl_get_mesh(self.h, *[m[k].ctypes.data for k in
                                                  ("x1", "x2", "x3", "Jac", "Jac2", "Ab", "At", "Ap")]))
        return m

    # -- blocks (reference order)
    def begin_step(self):
        self._chk(self.lib.tbl_begin_step(self.h))

    def recycle_inflow(self):
        self._chk(self.lib.tbl_recycle_inflow(self.h))

    def momentum_rhs(self):
        self._chk(self.lib.tbl_momentum_rhs(self.h))

    def diffusion_solve(self, errorstop=1.0e-6, mmax=1000):
        sw = (C.c_int * 3)()
        res = (C.c_double * 3)()
        self._chk(self.lib.tbl_diffusion_solve(self.h, errorstop, mmax, sw, res))
        return np.array(list(sw)), np.array(list(res))

    def ppe_rhs(self):
        self._chk(self.lib.tbl_ppe_rhs(self.h))

    def ppe_solve(self, errtol=1.0e-5, kmax=None, mmax=1000):
        kmax = self.p.kmax if kmax is None else kmax
        cy, it, tr = C.c_int(), C.c_int(), C.c_double()
        self._chk(self.lib.tbl_ppe_solve(self.h, errtol, kmax, mmax, C.byref(cy), C.byref(it), C.byref(tr)))
        return dict(cycles=cy.value, iters=it.value, true_res=tr.value)

    def project(self):
        self._chk(self.lib.tbl_project(self.h))

    def apply_bcs(self):
        self._chk(self.lib.tbl_apply_bcs(self.h))

    def div_check(self):
        d, s = C.c_double(), C.c_double()
        self._chk(self.lib.tbl_div_check(self.h, C.byref(d), C.byref(s)))
        return d.value, s.value

    def mass_flux(self):
        m = (C.c_double * 5)()
        self._chk(self.lib.tbl_mass_flux(self.h, m))
        return np.array(list(m))

    def stats_accumulate(self):
        self._chk(self.lib.tbl_stats_accumulate(self.h))

    def rescale(self):
        self._chk(self.lib.tbl_rescale(self.h))

    def ke_guard(self):
        ke, gd = C.c_double(), C.c_int()
        self._chk(self.lib.tbl_ke_guard(self.h, C.byref(ke), C.byref(gd)))
        return ke.value, gd.value

    def wall_shear(self):
        tau = np.zeros(self.nX)
        self._chk(self.lib.tbl_wall_shear(self.h, tau.ctypes.data))
        return tau

    def wall_shear_mean(self):
        """last completed wall-shear window: (tau_mean (nX), C_f (nX), windows done)"""
        tm, cf, nw = np.zeros(self.nX), np.zeros(self.nX), C.c_int(0)
        self._chk(self.lib.tbl_wall_shear_mean(self.h, tm.ctypes.data, cf.ctypes.data, C.byref(nw)))
        return tm, cf, nw.value

    def div_check_noalloc(self, ut_dev_ptr, out_dev_ptr):
        """tbl_div_check_noalloc on caller-owned device buffers (raw device
        addresses); stream-ordered, the caller synchronises before reading"""
        self._chk(self.lib.tbl_div_check_noalloc(self.h, C.c_void_p(ut_dev_ptr), C.c_void_p(out_dev_ptr)))

    def step(self, n=1):
        d = (Diag * max(n, 1))()
        self._chk(self.lib.tbl_step(self.h, n, d))
        return [d[q].as_dict() for q in range(n)]

    def apply_operator(self, op, x):
        out = self.zeros()
        self._chk(self.lib.tbl_apply_operator(self.h, op, _ptr(_f64(x)), _ptr(out)))
        return out

    # -- timers
    def timer_enable(self, names):
        mask = 0
        for n in names:
            mask |= 1 << KCLASSES.index(n)
        self._chk(self.lib.tbl_timer_enable(self.h, mask))

    def timer_read(self):
        cnt = (C.c_int * len(KCLASSES))()
        ms = (C.c_double * len(KCLASSES))()
        by = (C.c_double * len(KCLASSES))()
        self._chk(self.lib.tbl_timer_read(self.h, cnt, ms, by))
        return {n: dict(count=cnt[q], ms=ms[q], bytes=by[q]) for q, n in enumerate(KCLASSES) if cnt[q]}

    # -- discrete adjoint
    def adjoint_window(self, nsteps, i0, i1, wall_w=None, nsnap=10, forward_only=False, branch_mode=0):
        """Objective J (time/span-mean wall shear over 1-based stations i0..i1
        after each of `nsteps` steps) and, unless forward_only, dJ/d wall_w
        (nX, nY) by the discrete adjoint with binomial checkpointing.
        branch_mode 1 records the window's Peclet-switch pattern, 2 replays it
        (forward only).  Returns (J, grad or None, forward steps run, adjoint
        steps run)."""
        o = AdjointOpts(nsteps, nsnap, i0, i1, 1 if forward_only else 0, branch_mode)
        w = None if wall_w is None else _f64(wall_w)
        g = None if forward_only else np.zeros((self.nX, self.nY), order="F")
        J = C.c_double(0.0)
        nf, na = C.c_int(0), C.c_int(0)
        self._chk(self.lib.tbl_adjoint_window(self.h, C.byref(o), _ptr(w), C.byref(J), _ptr(g), C.byref(nf),
                                              C.byref(na)))
        return J.value, g, nf.value, na.value
