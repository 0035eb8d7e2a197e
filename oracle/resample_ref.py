"""CPU reference of spectral resampling between two cell-centred grids of one
periodic box (test infrastructure only, like the rest of oracle/).

No counterpart in the reference.  The checker evaluates the trigonometric
interpolant of the source samples directly at the destination's cell centres:
per axis, B[m, j] = (1/ns) sum_k c_k exp(2 pi i k (x_m - y_j) / L) over
k = -ns/2 .. ns/2 (c = 1/2 at the two ends for even ns; |k| <= (ns-1)/2 for
odd ns), x the destination centres and y the source centres, and the 3-D map
is B_x (x) B_y (x) B_z.  No FFT and no folding of modes is involved, unlike the
product path's gather tables."""
from __future__ import annotations

import numpy as np


def centres(n: int, L: float) -> np.ndarray:
    return -0.5 * L + (np.arange(n) + 0.5) * (L / n)


def interp_matrix(ns: int, nd: int, L: float) -> np.ndarray:
    if ns % 2 == 0:
        k = np.arange(-ns // 2, ns // 2 + 1)
        c = np.ones(k.size)
        c[0] = c[-1] = 0.5
    else:
        k = np.arange(-(ns - 1) // 2, (ns - 1) // 2 + 1)
        c = np.ones(k.size)
    x, y = centres(nd, L), centres(ns, L)
    ph = np.exp(2j * np.pi * k[None, None, :] * (x[:, None, None] - y[None, :, None]) / L)
    return (ph * c).sum(axis=2) / ns


def resample(a: np.ndarray, L, shape) -> np.ndarray:
    """a: (nx, ny, nz[, nc]) samples on the box of lengths L (scalar or 3); returns the
    interpolant at the centres of the grid `shape`."""
    a = np.asarray(a, dtype=complex)
    if a.ndim == 3:
        a = a[..., None]
    Ls = [float(L)] * 3 if np.isscalar(L) else [float(v) for v in L]
    B = [interp_matrix(a.shape[d], shape[d], Ls[d]) for d in range(3)]
    out = np.einsum("ai,bj,ck,ijkn->abcn", B[0], B[1], B[2], a, optimize=True)
    return np.asfortranarray(out)
