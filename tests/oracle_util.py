"""Helpers shared by the oracle pin tests (dense assembly on tiny grids)."""
import numpy as np

from oracle import orc


def small_grid(nX=10, nY=8, nZ=10, P=1, **kw):
    kw.setdefault("i_rec", max(3, int(round(0.8 * (nX - 1))) + 1 - 1))
    kw.setdefault("planes", tuple(max(2, min(nX - 1, int(round(p / 90.0 * nX)))) for p in orc.REF_PLANES))
    return orc.Grid(nX, nY, nZ, P=P, **kw)


def interior_index(g):
    """Linear (Fortran) offsets of interior nodes in lexicographic (i fastest) order."""
    i = np.arange(1, g.nX - 1)
    j = np.arange(1, g.nY - 1)
    k = np.arange(1, g.nZ - 1)
    K, J, I = np.meshgrid(k, j, i, indexing="ij")
    return (I + g.nX * (J + g.nY * K)).ravel()


def ijk_of(g, lin):
    i = lin % g.nX
    j = (lin // g.nX) % g.nY
    k = lin // (g.nX * g.nY)
    return i + 1, j + 1, k + 1


def dense_operator(g, apply):
    """Dense matrix of x -> apply(x) restricted to interior rows/cols (boundary of x = 0)."""
    idx = interior_index(g)
    n = len(idx)
    A = np.zeros((n, n))
    for c in range(n):
        x = g.zeros()
        x.reshape(-1, order="F")[idx[c]] = 1.0
        A[:, c] = apply(x).reshape(-1, order="F")[idx]
    return A, idx


def to_interior(g, field, idx=None):
    idx = interior_index(g) if idx is None else idx
    return np.asarray(field).reshape(-1, order="F")[idx]


def from_interior(g, vec, idx=None, base=None):
    idx = interior_index(g) if idx is None else idx
    out = g.zeros() if base is None else base.copy(order="F")
    out.reshape(-1, order="F")[idx] = vec
    return out
