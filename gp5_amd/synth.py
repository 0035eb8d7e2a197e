"""Seeded synthetic initial states of the benchmark workloads.

One counter-based stream: u_i = splitmix64(seed * 2^40 + i) mapped to
[-1, 1) from its top 53 bits.  The same stream is restated in the oracle's C
code, so CPU and GPU runs start from bit-identical inputs.  Arrays are
complex128 in the reference layout (Nx, Ny, Nz, nC), Fortran order.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """splitmix64 counter stream mapped to [-1, 1)."""
    with np.errstate(over="ignore"):
        x = (np.uint64(seed) << np.uint64(40)) + np.uint64(start) + np.arange(count, dtype=np.uint64)
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return 2.0 * ((x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def axes(shape: Sequence[int], left: Sequence[float], right: Sequence[float]):
    """Cell-centred coordinates x_i = left + (i + 1/2) dx (reference src/geometry.cpp:165-188)."""
    return [left[d] + (np.arange(shape[d]) + 0.5) * ((right[d] - left[d]) / shape[d]) for d in range(3)]


def r2_grid(shape, left, right, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    x, y, z = axes(shape, left, right)
    return ((x - center[0])[:, None, None] ** 2 + (y - center[1])[None, :, None] ** 2
            + (z - center[2])[None, None, :] ** 2)


def dV(shape, left, right) -> float:
    v = 1.0
    for d in range(3):
        v *= (right[d] - left[d]) / shape[d]
    return v


def gaussian_noisy(shape, left, right, sigma: float, noise: float, seed: int, N: float) -> np.ndarray:
    """exp(-r^2/(2 sigma^2)) + noise*(u1 + i u2), normalised to N (harmonic-trap workload)."""
    M = int(np.prod(shape))
    g = np.exp(-r2_grid(shape, left, right) / (2 * sigma * sigma))
    u = uniform(seed, 2 * M)
    re = u[0::2].reshape(shape, order="F")
    im = u[1::2].reshape(shape, order="F")
    psi = (g + noise * re) + 1j * (noise * im)
    psi *= np.sqrt(N / (np.sum(np.abs(psi) ** 2) * dV(shape, left, right)))
    return np.asfortranarray(psi[..., None])


def flat_top_droplet(shape, left, right, radius: float, width: float, phase_noise: float, seed: int,
                     N: float) -> np.ndarray:
    """A * s(r) * exp(i * phase_noise * u), s = (1 - tanh((r - R)/w))/2, A fixed by N (LHY droplet workload)."""
    M = int(np.prod(shape))
    r = np.sqrt(r2_grid(shape, left, right))
    s = 0.5 * (1.0 - np.tanh((r - radius) / width))
    A = np.sqrt(N / (np.sum(s * s) * dV(shape, left, right)))
    u = uniform(seed, M).reshape(shape, order="F")
    psi = A * s * np.exp(1j * phase_noise * u)
    return np.asfortranarray(psi[..., None])


def harmonic_potential(shape, left, right, omegas: Sequence[Sequence[float]]) -> np.ndarray:
    """V_c = 1/2 sum_d omega_{c,d}^2 x_d^2 (reference src/externalPotential.cpp:11-32), real FP64."""
    x, y, z = axes(shape, left, right)
    out = np.empty(tuple(shape) + (len(omegas),), order="F")
    for c, w in enumerate(omegas):
        out[..., c] = 0.5 * (w[0] ** 2 * x[:, None, None] ** 2 + w[1] ** 2 * y[None, :, None] ** 2
                             + w[2] ** 2 * z[None, None, :] ** 2)
    return out


# Workload definitions ------------------------------------------------------

def droplet_256(n: int = 256) -> Tuple[Tuple[int, int, int], Tuple[float, ...], Tuple[float, ...], np.ndarray]:
    """Untrapped LHY droplet, n^3 periodic box L = 40, flat top r = 6 (width 0.5) scaled
    to N = 3000 with seeded (seed 2) 1e-4 phase noise."""
    shape = (n, n, n)
    left, right = (-20.0,) * 3, (20.0,) * 3
    return shape, left, right, flat_top_droplet(shape, left, right, 6.0, 0.5, 1e-4, 2, 3000.0)


def droplet_collision(n: int = 512, L: float = 64.0, N: float = 4000.0, x0: float = 8.0, k: float = 0.5):
    """Two untrapped LHY droplets of N atoms each centred at x = -x0 and x = +x0 of the
    periodic box [-L/2, L/2]^3 on an n^3 grid, moving toward each other along x:
    psi = A [s(r-) e^{+i q x} + s(r+) e^{-i q x}], s a flat top (width 0.5) of the
    saturation-density radius (3N / 4 pi)^{1/3}, q = k snapped to the nearest 2 pi m / L,
    A fixed by the total norm 2N.  Built slab by slab along z (2 GiB at 512^3).
    Returns (shape, left, right, psi (n, n, n, 1) complex128 Fortran order, q)."""
    shape = (n, n, n)
    left, right = (-0.5 * L,) * 3, (0.5 * L,) * 3
    q = 2.0 * np.pi / L * round(k * L / (2.0 * np.pi))
    R = droplet_radius(N)
    x = axes(shape, left, right)[0]
    psi = np.empty((n, n, n, 1), dtype=np.complex128, order="F")
    ex_p, ex_m = np.exp(1j * q * x), np.exp(-1j * q * x)
    yy = x[None, :] ** 2
    for kz in range(n):
        zz = x[kz] ** 2
        rm = np.sqrt((x[:, None] + x0) ** 2 + yy + zz)
        rp = np.sqrt((x[:, None] - x0) ** 2 + yy + zz)
        sm = 0.5 * (1.0 - np.tanh((rm - R) / 0.5))
        sp = 0.5 * (1.0 - np.tanh((rp - R) / 0.5))
        psi[:, :, kz, 0] = sm * ex_p[:, None] + sp * ex_m[:, None]
    norm = 0.0
    for kz in range(n):
        norm += float(np.sum(np.abs(psi[:, :, kz, 0]) ** 2))
    psi *= np.sqrt(2.0 * N / (norm * dV(shape, left, right)))
    return shape, left, right, psi, q


def harmonic_64(n: int = 64):
    """Single-component GP, n^3 on [-10, 10]^3, Gaussian sigma = 1.5 plus seeded (seed 1)
    1e-3 complex uniform noise, normalised to N = 1; harmonic trap omega = (1,1,1)."""
    shape = (n, n, n)
    left, right = (-10.0,) * 3, (10.0,) * 3
    psi = gaussian_noisy(shape, left, right, 1.5, 1e-3, 1, 1.0)
    V = harmonic_potential(shape, left, right, [(1.0, 1.0, 1.0)])
    return shape, left, right, psi, V


# ---------------------------------------------------------------------------
# batches of droplet ground states
# ---------------------------------------------------------------------------
BATCH_GRIDS = (64, 96, 128, 192)
BATCH_DX = 0.35


def droplet_radius(N: float) -> float:
    """Radius of an LHY droplet of N atoms at the saturation density 1 of
    -3 rho + 2.5 rho^{3/2} (zero pressure: rho = 1)."""
    return (3.0 * N / (4.0 * np.pi)) ** (1.0 / 3.0)


def batch_grid(N: float, dx: float = BATCH_DX, grids: Sequence[int] = BATCH_GRIDS) -> int:
    """Smallest grid whose box n*dx spans six droplet radii (the largest grid otherwise)."""
    R = droplet_radius(N)
    for n in grids:
        if n * dx >= 6.0 * R:
            return int(n)
    return int(grids[-1])


def droplet_batch(count: int = 64, seed: int = 3, Nmin: float = 50.0, Nmax: float = 5000.0,
                  dx: float = BATCH_DX):
    """Instances of the batched ground-state workload: N log-uniform in [Nmin, Nmax]
    from the seeded stream, grid n from the predicted radius, box L = n dx centred on
    the origin, Gaussian seed of width 0.6 R (amplitude set by the first renormalisation).
    Returns (n, L, N, psi0) lists."""
    u = 0.5 * (uniform(seed, count) + 1.0)
    N = np.exp(np.log(Nmin) + u * (np.log(Nmax) - np.log(Nmin)))
    n = [batch_grid(v, dx) for v in N]
    L = [k * dx for k in n]
    psi0 = []
    for i in range(count):
        x = -0.5 * L[i] + (np.arange(n[i]) + 0.5) * dx
        r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
        s = 0.6 * droplet_radius(N[i])
        psi0.append(np.asfortranarray(np.exp(-r2 / (2.0 * s * s)).astype(np.complex128)))
    return n, L, [float(v) for v in N], psi0
