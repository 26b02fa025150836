"""FP64 CPU reference adjoint of the TBL step on tiny grids (test infrastructure).

The step of an adjoint window (dnssetbl/dnssetbl.f90:253-797 with the
recycling inflow frozen and both linear systems solved exactly) is composed
from the oracle's block functions:

* recycling inflow, momentum RHS, PPE right-hand side, projection and
  boundary conditions come from the oracle (oracle/tbl_oracle.c);
* the Helmholtz solve is u_int = H^{-1} (Q + h B u_faces) with H = I - hA;
* the Poisson solve is phi_int = A^{-1} b.

Both use dense LU factors of the interior Poisson operator A. A is assembled
by probing the oracle's `poisson_apply`, and B u_faces is the same apply on a
field that holds only the face values.

The gradient of J = (1/n) sum_n mean_{i0..i1} tau_n / (nY-2) is accumulated
in reverse:

    lam_{n-1} = (dS_n/dx)^T lam_n + dJ_{n-1}/dx,  g += (dS_n/dw)^T lam_n.

Each entry of the vector-Jacobian product lam^T dS/dx_e is <lam, S(x + eps e)
- S(x - eps e)> / (2 eps). Between branch switches the step is
piecewise quadratic in (Ut, Press, wall_w): the Kawamura-Kuwahara |u| and
the Peclet switch are fixed, and the bottom and inflow pressure rows are
quadratic. Central differences are therefore exact there up to rounding.
"""
import numpy as np
from scipy.linalg import lu_factor, lu_solve

from oracle import orc
from tests.oracle_util import dense_operator


class ExactStep:
    """(Ut, Press, wall_w, nstep) -> (Ut, Press) of one window step, exact solves"""

    def __init__(self, g, st):
        self.g, self.st = g, st
        A, idx = dense_operator(g, g.poisson_apply)
        self.idx = idx
        self.h = g.dt / g.Re
        self.A_lu = lu_factor(A)
        self.H_lu = lu_factor(np.eye(len(idx)) - self.h * A)

    def __call__(self, U, P, vw, nstep):
        g, st, idx = self.g, self.st, self.idx
        Ut = np.array(U, order="F", copy=True)
        Press = np.array(P, order="F", copy=True)
        Utold = Ut.copy(order="F")
        g.recycle_inflow(Ut, nstep, st.lamda, st.avg, st.Wt, st.newk, st.Ume, st.Ufluc_rec)
        _, Q = g.momentum_rhs(Ut, Press)
        for p in range(3):
            u = np.array(Ut[..., p], order="F")
            faces = u.copy(order="F")
            faces.reshape(-1, order="F")[idx] = 0.0
            bnd = g.poisson_apply(faces).reshape(-1, order="F")[idx]
            rhs = np.asarray(Q[..., p]).reshape(-1, order="F")[idx] + self.h * bnd
            u.reshape(-1, order="F")[idx] = lu_solve(self.H_lu, rhs)
            Ut[..., p] = u
        pRhs = g.ppe_rhs(Ut)
        phi = g.zeros()
        phi.reshape(-1, order="F")[idx] = lu_solve(self.A_lu, np.asarray(pRhs).reshape(-1, order="F")[idx])
        g.project(phi, Press, Ut)
        g.apply_bcs(Ut, Utold, Press, vw)
        return Ut, Press


def objective_weights(g, nsteps, i0, i1):
    """dJ_n/dUt (the same every step: tau is linear in Ut), shape (nX,nY,nZ,3)"""
    cJ = 1.0 / (nsteps * (i1 - i0 + 1) * (g.nY - 2))
    base = g.zeros(3)
    w = g.zeros(3)
    flat = w.reshape(-1, order="F")
    bflat = base.reshape(-1, order="F")
    # tau depends on the wall-adjacent planes k = 1..3 only
    nX, nY = g.nX, g.nY
    for p in (0, 2):
        for k in range(3):
            for j in range(nY):
                for i in range(nX):
                    e = i + nX * (j + nY * (k + g.nZ * p))
                    bflat[e] = 1.0
                    tp = g.wall_shear(base)
                    bflat[e] = 0.0
                    flat[e] = cJ * tp[i0 - 1:i1].sum()
    return w


def objective(g, states, i0, i1):
    n = len(states) - 1
    cJ = 1.0 / (n * (i1 - i0 + 1) * (g.nY - 2))
    return sum(cJ * g.wall_shear(U)[i0 - 1:i1].sum() for U, _ in states[1:])


def forward(step, U0, P0, vw, nstep0, nsteps):
    states = [(np.array(U0, order="F"), np.array(P0, order="F"))]
    for n in range(1, nsteps + 1):
        states.append(step(*states[-1], vw, nstep0 + n))
    return states


def reference_adjoint(step, U0, P0, vw, nstep0, nsteps, i0, i1, eps=1e-3):
    """(J, dJ/dwall_w) by reverse accumulation of per-step vector-Jacobian products"""
    g = step.g
    states = forward(step, U0, P0, vw, nstep0, nsteps)
    J = objective(g, states, i0, i1)
    wJ = objective_weights(g, nsteps, i0, i1)
    lamU, lamP = wJ.copy(order="F"), g.zeros()
    grad = np.zeros((g.nX, g.nY), order="F")

    def vjp(arr, e, U, P, n):
        # <lam, dS/dx_e>: the two perturbed states are differenced elementwise
        # first, so the large state values cancel before the dot product
        v0 = arr[e]
        arr[e] = v0 + eps
        Up, Pp = step(U, P, vw, nstep0 + n)
        arr[e] = v0 - eps
        Um, Pm = step(U, P, vw, nstep0 + n)
        arr[e] = v0
        return float(np.sum(lamU * (Up - Um)) + np.sum(lamP * (Pp - Pm))) / (2 * eps)

    for n in range(nsteps, 0, -1):
        U, P = states[n - 1]
        U, P = U.copy(order="F"), P.copy(order="F")
        newU, newP = g.zeros(3), g.zeros()
        for arr, out in ((U, newU), (P, newP)):
            fa, fo = arr.reshape(-1, order="F"), out.reshape(-1, order="F")
            for e in range(fa.size):
                fo[e] = vjp(fa, e, U, P, n)
        vflat, gflat = vw.reshape(-1, order="F"), grad.reshape(-1, order="F")
        for e in range(vflat.size):
            gflat[e] += vjp(vflat, e, U, P, n)
        lamU, lamP = newU, newP
        if n - 1 >= 1:
            lamU = lamU + wJ
    return J, grad
