"""Two-level ground states of batched droplets: converge on a grid of half the
resolution over the same box, carry each state to the full grid by spectral
resampling on the GPU (gp5_batch_resample), and finish there.

Why it pays: imaginary time removes the error of a start at the rate of the
slowest mode the error contains (for droplets, the breathing mode), so the
number of fine-grid steps grows with the decades of energy error the start has
to lose.  A ground state of the half-resolution grid differs from the
full-grid one only by what the coarse grid cannot resolve, about 1e-4 in
amplitude and 1e-7 in relative energy for the 0.35-spaced droplet grids, so the
fine grid converges in two or three checks.  The coarse grid costs an eighth
per step and allows four times the imaginary-time step (dt = dt_factor dx^2).

The coarse level itself starts from Gaussians or from the learned warm start
(gp5_amd.surrogate), evaluated on the coarse grid.  No counterpart in the
reference, which evolves one field from a file.
"""
from __future__ import annotations

import time
from typing import Callable, Optional, Sequence

import numpy as np

from . import capi


def coarse_size(n: int, minimum: int = 16) -> int:
    """Half the grid (same box, twice the spacing) when n is even and the half is
    at least `minimum`; otherwise n."""
    return n // 2 if n % 2 == 0 and n // 2 >= minimum else n


def gaussian_seed(n: int, L: float, N: float) -> np.ndarray:
    """Spherical Gaussian of width 0.6 R, R = (3N / 4 pi)^{1/3} (saturation density 1)."""
    x = -0.5 * L + (np.arange(n) + 0.5) * (L / n)
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    s = 0.6 * (3.0 * N / (4.0 * np.pi)) ** (1.0 / 3.0)
    return np.asfortranarray(np.exp(-r2 / (2.0 * s * s)).astype(np.complex128))


def two_level(n: Sequence[int], L: Sequence[float], N: Sequence[float], f: capi.FunctionalDesc, *,
              seed: Optional[Callable[[int, int, float, float], np.ndarray]] = None, dt_factor: float = 0.1,
              tol: float = 1e-10, coarse_tol: float = 1e-9, check_every: int = 100, max_steps: int = 200000,
              tune: bool = True, potentials: Optional[Callable[[int, int, float], np.ndarray]] = None):
    """Ground states of instances (n[i]^3 grid over [-L[i]/2, L[i]/2]^3, norm N[i]).

    seed(i, n, L, N) -> complex (n, n, n) start on the coarse grid (default: Gaussian);
    potentials(i, n, L) -> real (n, n, n) potential or None.  Returns (fine batch, report);
    the caller downloads the states from the batch and closes it."""
    count = len(n)
    nc = [coarse_size(int(k)) for k in n]
    report = {"coarse_grids": nc}
    t0 = time.perf_counter()
    cb = capi.Batch(nc, L, N)
    for i in range(count):
        a = seed(i, nc[i], L[i], N[i]) if seed else gaussian_seed(nc[i], L[i], N[i])
        cb.upload(i, a)
        if potentials:
            V = potentials(i, nc[i], L[i])
            if V is not None:
                cb.set_potential(i, V)
    report["seed_s"] = time.perf_counter() - t0
    report["tune_s"] = 0.0
    if tune:
        t = time.perf_counter()
        cb.tune(f, dt_factor, 2)
        report["tune_s"] += time.perf_counter() - t
    steps_c, en_c, ms_c = cb.ground_state(f, dt_factor, coarse_tol, check_every, max_steps)
    report.update(coarse_s=ms_c / 1e3, coarse_steps=steps_c, coarse_energy=en_c)
    t1 = time.perf_counter()
    fb = capi.Batch(n, L, N)
    for i in range(count):
        fb.resample_from(i, cb, i)
        if potentials:
            V = potentials(i, int(n[i]), L[i])
            if V is not None:
                fb.set_potential(i, V)
    cb.close()
    report["resample_s"] = time.perf_counter() - t1
    if tune:
        t = time.perf_counter()
        fb.tune(f, dt_factor, 2)
        report["tune_s"] += time.perf_counter() - t
    steps_f, en_f, ms_f = fb.ground_state(f, dt_factor, tol, check_every, max_steps)
    report.update(fine_s=ms_f / 1e3, steps=steps_f, energy=en_f)
    # tuning is reported apart, as for single-level runs
    report["total_s"] = report["seed_s"] + report["coarse_s"] + report["resample_s"] + report["fine_s"]
    return fb, report
