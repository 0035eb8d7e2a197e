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
    # spherical: the transpose is Fortran-ordered with the same values up to the
    # summation order of r2, which saves the upload a transposing copy
    return np.exp(-r2 / (2.0 * s * s)).astype(np.complex128).T


def level_sizes(n: int, levels: int, minimum: int = 16, L: Optional[float] = None,
                max_spacing: Optional[float] = None):
    """Grids from the coarsest to n, halving while the grid is even, the half is at
    least `minimum` and (with L and max_spacing) its spacing L / size stays at most
    max_spacing; fewer levels when it stops."""
    sizes = [int(n)]
    for _ in range(levels - 1):
        h = coarse_size(sizes[0], minimum)
        if h == sizes[0] or (max_spacing is not None and L is not None and L / h > max_spacing):
            break
        sizes.insert(0, h)
    return sizes


def multilevel(n: Sequence[int], L: Sequence[float], N: Sequence[float], f: capi.FunctionalDesc, *,
               levels: int = 2, seed: Optional[Callable[[int, int, float, float], np.ndarray]] = None,
               dt_factor: float = 0.1, tol: float = 1e-10, coarse_tol: float = 1e-9, check_every: int = 100,
               max_steps: int = 200000, tune: bool = True, max_spacing: Optional[float] = None,
               potentials: Optional[Callable[[int, int, float], np.ndarray]] = None,
               seed_many: Optional[Callable[[list], list]] = None):
    """Ground states of instances (n[i]^3 grid over [-L[i]/2, L[i]/2]^3, norm N[i]),
    converged level by level on grids of the same box, each level halving the
    resolution of the next (level_sizes); every level but the last stops at
    coarse_tol, the last at tol.  max_spacing (box units) bounds the spacing of the
    coarse grids.

    Two levels are the measured optimum on the droplet batch (0.35 spacing): a third
    level (n/4) left the n/2 level at the same step count and cost more in total, and
    grids of spacing 1.4-2.8 let some droplets settle on a higher-energy plateau
    that the energy-change rule took for convergence.

    seed(i, g, L, N) -> complex (g, g, g) start on the coarsest grid g (default:
    Gaussian), or seed_many([(i, g, L, N), ...]) -> the starts of all instances in
    order (any iterable; timed the same way); potentials(i, g, L) -> real (g, g, g)
    potential or None.  Returns
    (finest batch, report); the caller downloads the states from the batch and closes it."""
    count = len(n)
    sizes = [level_sizes(int(k), levels, L=float(Lk), max_spacing=max_spacing) for k, Lk in zip(n, L)]
    depth = max(len(z) for z in sizes)
    # pad instances with fewer levels by repeating their coarsest grid
    grids = [[z[0]] * (depth - len(z)) + z for z in sizes]
    report = {"grids": grids, "coarse_grids": [g[-2] if depth > 1 else g[-1] for g in grids], "levels": [],
              "tune_s": 0.0, "resample_s": 0.0}

    def load_potentials(batch, lev):
        if potentials:
            for i in range(count):
                V = potentials(i, grids[i][lev], L[i])
                if V is not None:
                    batch.set_potential(i, V)

    t0 = time.perf_counter()
    b = capi.Batch([g[0] for g in grids], L, N)
    nb = None
    try:
        if seed_many:
            for i, psi in enumerate(seed_many([(i, grids[i][0], L[i], N[i]) for i in range(count)])):
                b.upload(i, psi)
        else:
            for i in range(count):
                g = grids[i][0]
                b.upload(i, seed(i, g, L[i], N[i]) if seed else gaussian_seed(g, L[i], N[i]))
        load_potentials(b, 0)
        report["seed_s"] = time.perf_counter() - t0
        for lev in range(depth):
            if lev > 0:
                t = time.perf_counter()
                nb = capi.Batch([g[lev] for g in grids], L, N)
                for i in range(count):
                    nb.resample_from(i, b, i)
                load_potentials(nb, lev)
                b.close()
                b, nb = nb, None
                report["resample_s"] += time.perf_counter() - t
            if tune:
                t = time.perf_counter()
                b.tune(f, dt_factor, 2)
                report["tune_s"] += time.perf_counter() - t
            last = lev == depth - 1
            steps, en, ms = b.ground_state(f, dt_factor, tol if last else coarse_tol, check_every, max_steps)
            report["levels"].append({"grids": sorted(set(g[lev] for g in grids)), "seconds": ms / 1e3,
                                     "steps": steps, "energy": en})
    except BaseException:
        # release the device batches now rather than when the traceback is dropped
        for x in (nb, b):
            if x is not None:
                try:
                    x.close()
                except Exception:
                    pass
        raise
    fin = report["levels"][-1]
    report["coarse_s"] = sum(lv["seconds"] for lv in report["levels"][:-1])
    report["coarse_steps"] = report["levels"][-2]["steps"] if depth > 1 else fin["steps"]
    report.update(fine_s=fin["seconds"], steps=fin["steps"], energy=fin["energy"])
    # tuning is reported apart, as for single-level runs
    report["total_s"] = report["seed_s"] + report["coarse_s"] + report["resample_s"] + report["fine_s"]
    return b, report


def two_level(n, L, N, f, **kw):
    """multilevel with one coarse level (half the resolution)."""
    return multilevel(n, L, N, f, levels=2, **kw)


def droplet_pair(n: int = 512, L: float = 64.0, N: float = 4000.0, x0: float = 8.0, k: float = 0.5,
                 b: float = 2.5, tol: float = 1e-10, use_surrogate: bool = True, dt_factor: float = 0.1,
                 check_every: int = 100):
    """Initial state of the large collision workload: two LHY droplets of N atoms built
    from one ground state of the periodic box [-L/2, L/2]^3 on n^3 (dx = L/n), computed
    by two_level (coarse grid n/2, seeded by the learned warm start when its weights
    are present, else by a Gaussian; FP64 imaginary time to `tol` relative energy
    change per `check_every` steps), moved to x = -x0 and x = +x0 and given the
    momenta +q and -q toward each other:
        psi = g(x + x0, y, z) e^{+i q x} + g(x - x0, y, z) e^{-i q x},
    q = k snapped to 2 pi m / L (periodic), x0 a whole number of cells (an exact
    shift of the samples), the sum normalised to 2N.
    Returns (shape, left, right, psi (n, n, n, 1) complex128 Fortran order, q, info)."""
    import os

    dx = L / n
    shift = int(round(x0 / dx))
    if abs(shift * dx - x0) > 1e-12 * L:
        raise ValueError(f"x0 = {x0} is not a whole number of cells of {dx}")
    q = 2.0 * np.pi / L * round(k * L / (2.0 * np.pi))
    f = capi.functional_desc(capi.GP5_FUNC_LHY, lhy_a=-3.0, lhy_b=b)
    kw = {}
    start = "gaussian"
    if use_surrogate:
        from . import surrogate

        if os.path.exists(surrogate.DEFAULT_WEIGHTS):
            sur = surrogate.Surrogate()
            kw["seed"] = lambda i, g, Lg, Ng: sur.seed(Ng, (0.0, 0.0, 0.0), b, g, Lg)
            start = "learned warm start"
    t0 = time.perf_counter()
    fb, rep = two_level([n], [L], [N], f, dt_factor=dt_factor, tol=tol, check_every=check_every, **kw)
    try:
        g = fb.download(0)
    finally:
        fb.close()
    seconds = time.perf_counter() - t0
    x = -0.5 * L + (np.arange(n) + 0.5) * dx
    ep, em = np.exp(1j * q * x), np.exp(-1j * q * x)
    psi = np.empty((n, n, n, 1), dtype=np.complex128, order="F")
    # g(x + x0) at sample i is g's sample i + shift (periodic), g(x - x0) the sample i - shift
    psi[..., 0] = np.roll(g, -shift, axis=0) * ep[:, None, None] + np.roll(g, shift, axis=0) * em[:, None, None]
    del g
    norm = 0.0
    for kz in range(n):
        norm += float(np.sum(np.abs(psi[:, :, kz, 0]) ** 2))
    psi *= np.sqrt(2.0 * N / (norm * dx ** 3))
    info = {"ground_state": f"one N={N:g} droplet on {n}^3 (L={L:g}), two-level from {rep['coarse_grids'][0]}^3, "
                            f"start: {start}, FP64 imaginary-time RK4 to {tol:g}",
            "seconds": seconds, "coarse_steps": int(np.abs(rep["coarse_steps"][0])),
            "fine_steps": int(np.abs(rep["steps"][0])), "converged": bool(rep["steps"][0] > 0),
            "energy_per_atom": float(rep["energy"][0] / N), "overlap_norm_factor": float(2.0 * N / (norm * dx ** 3))}
    return (n, n, n), (-0.5 * L,) * 3, (0.5 * L,) * 3, psi, q, info


def seed_error_curve(per_class: int = 4, grids=(64, 96), eps=(0.3, 0.1, 0.03, 0.01, 3e-3, 1e-3, 3e-4, 1e-4, 1e-5),
                     dt_factor: float = 0.1, tol: float = 1e-10, check_every: int = 100):
    """Steps to convergence against the error of the start.  For droplets of the
    batched workload (per_class of each grid in `grids`), a tight ground state g
    (1e-12) is computed cold; starts s = (1 - e) g + e c (c: the cold Gaussian seed),
    renormalised, span relative L2 errors from ~e down to 1e-5 of g.  All starts, the
    cold seeds and the network's seeds run in one batch to `tol`.  Returns per start:
    the median relative L2 error |s - g| / |g|, the total steps and the speed-up in
    steps over the cold starts."""
    import os

    from . import synth

    n0, L0, N0, psi0 = synth.droplet_batch(64, 3)
    pick = []
    for g in grids:
        pick += [i for i in range(len(n0)) if n0[i] == g][:per_class]
    n, L, N = [n0[i] for i in pick], [L0[i] for i in pick], [N0[i] for i in pick]
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    b = capi.Batch(n, L, N)
    try:
        for i, k in enumerate(pick):
            b.upload(i, psi0[k])
        b.ground_state(f, dt_factor, 1e-12, check_every, 200000)
        gs = [b.download(i) for i in range(len(n))]
    finally:
        b.close()

    def unit(a, Nk, h):
        return a * np.sqrt(Nk / (np.sum(np.abs(a) ** 2) * h ** 3))

    starts = {"cold": [], "network": []}
    for e in eps:
        starts[f"{e:g}"] = []
    sur = None
    from . import surrogate

    if os.path.exists(surrogate.DEFAULT_WEIGHTS):
        sur = surrogate.Surrogate()
    for i, k in enumerate(pick):
        h = L[i] / n[i]
        c = unit(psi0[k], N[i], h)
        starts["cold"].append(c)
        for e in eps:
            starts[f"{e:g}"].append(unit((1 - e) * gs[i] + e * c, N[i], h))
        if sur is not None:
            starts["network"].append(sur.seed(N[i], (0.0, 0.0, 0.0), 2.5, n[i], L[i]))
    if sur is None:
        del starts["network"]
    names = list(starts)
    inst = [(name, i) for name in names for i in range(len(n))]
    bb = capi.Batch([n[i] for _, i in inst], [L[i] for _, i in inst], [N[i] for _, i in inst])
    try:
        for j, (name, i) in enumerate(inst):
            bb.upload(j, starts[name][i])
        steps, en, ms = bb.ground_state(f, dt_factor, tol, check_every, 200000)
    finally:
        bb.close()
    out = {"instances": [f"{n[i]}^3 N={N[i]:.0f}" for i in range(len(n))], "tol": tol, "rows": []}
    cold_steps = None
    for name in names:
        idx = [j for j, (nm, _) in enumerate(inst) if nm == name]
        err = [float(np.linalg.norm(unit(starts[name][i], N[i], L[i] / n[i]) - gs[i]) / np.linalg.norm(gs[i]))
               for _, i in (inst[j] for j in idx)]
        tot = int(np.abs(steps[idx]).sum())
        if name == "cold":
            cold_steps = tot
        out["rows"].append({"start": name, "rel_l2_median": float(np.median(err)), "rel_l2_max": float(np.max(err)),
                            "steps_total": tot, "converged": int((steps[idx] > 0).sum())})
    for r in out["rows"]:
        r["speedup_in_steps_vs_cold"] = cold_steps / r["steps_total"]
    return out


def _main(argv=None) -> int:
    """python -m gp5_amd.ground_state advance [--count 64] [--steps 10]: the batched
    droplet workload (synth.droplet_batch) advanced eagerly, class by class, after
    tuning; prints one JSON line with each class's time per step.  A fixed, short
    workload of the batch kernels for rocprofv3 (kernel traces and counters)."""
    import argparse
    import json

    from . import synth

    ap = argparse.ArgumentParser(prog="python -m gp5_amd.ground_state")
    ap.add_argument("what", choices=["advance", "curve"])
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--dt-factor", type=float, default=0.1)
    args = ap.parse_args(argv)
    if args.what == "curve":
        print(json.dumps(seed_error_curve(dt_factor=args.dt_factor)))
        return 0
    n, L, N, psi0 = synth.droplet_batch(args.count, 3)
    b = capi.Batch(n, L, N)
    try:
        for i, a in enumerate(psi0):
            b.upload(i, a)
        f = capi.functional_desc(capi.GP5_FUNC_LHY)
        report = b.tune(f, args.dt_factor, 2)
        b.advance(f, args.dt_factor, 2)  # warm-up
        t0 = time.perf_counter()
        b.advance(f, args.dt_factor, args.steps)
        wall = time.perf_counter() - t0
    finally:
        b.close()
    print(json.dumps({"workload": f"{args.count} droplets (seed 3) on " + ", ".join(
        f"{g}^3 x{sum(1 for k in n if k == g)}" for g in synth.BATCH_GRIDS) + f", {args.steps} eager RK4 steps",
                      "seconds": wall, "ms_per_step": 1e3 * wall / args.steps, "tuning": report.strip().splitlines()}))
    return 0


if __name__ == "__main__":
    import sys

    sys.exit(_main())
