#!/usr/bin/env python3
"""Headline benchmark: FP64 real-time RK4 steps/s of the LHY droplet at 256^3 on one MI355X.

Workload: untrapped LHY droplet (-3|psi|^2 psi + 2.5|psi|^3 psi), 256^3 periodic box
L = 40, complex128, flat-top initial profile of radius 6 scaled to N = 3000 with seeded
1e-4 phase noise, real-time RK4 dt = 5e-4 (the reference's RK4Stepper algebra,
src/stepper.cpp:46-70).  Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  `roofline` prices the dominant kernel against the 8 TB/s HBM
peak from its algorithmic bytes per launch and its average launch duration measured
here with HIP events on the context's stream; `cpu_baseline` times the CPU oracle
(the C restatement of the reference path) on this host's cores for a bounded sample.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--n 256] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 steps/s at 256³ FP64 + ground-state time, surrogate-seeded vs cold start"
HBM_PEAK_GBPS = 8000.0

# Algorithmic HBM bytes per launch of each kernel kind, in units of F = 16*M bytes
# (one complex128 field): an axis pass reads and writes the field once.  The x-inverse
# epilogue of RK4 stage s never stores the stage vector u_s: it rebuilds it from psi and
# the previous slope K, so stage 1 reads T, psi and writes K1 (3F); stage 2 reads T, psi,
# K1 and writes K2, S = K1 + 2K2 (5F); stage 3 reads T, psi, K2 and writes K3 (4F);
# stage 4 reads T, psi, K3, S and writes psi' (5F): 4.25F on average.  The fused variant
# also writes the x transform of the next stage's input (one more F: 5.25F on average).
KIND_BYTES_F = {
    "fft_x_forward": 2.0,
    "fft_y_pass": 2.0,
    "fft_z_fwd_mul_inv": 2.0,
    "fft_x_inverse_epilogue": 4.25,
    "fft_x_inv_epilogue_fwd": 5.25,
}
STEP_MODEL_F = 29.0  # ideal compulsory traffic of one RK4 step (fused, one r/w per FFT)


def _progress(msg: str) -> None:
    """Phase markers on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _affinity() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(psi0, shape, left, right, budget_s: float = 25.0):
    """Time the CPU oracle (restatement of GP5's FFTW path) on this host: 1 RK4 step of the
    same workload with OpenMP over the available cores (capped at 16)."""
    from oracle import oracle as O

    O.build()
    cores = min(_affinity(), 16)
    O.set_threads(cores)
    geo = O.Geometry(shape, left, right, 1)
    F = O.Functional(kind=O.OR_LHY)
    steps = 0
    t0 = time.perf_counter()
    psi = psi0
    while True:
        psi = O.advance(F, psi, geo, stepper="RK4", dt=5e-4, imaginary=False, nsteps=1)
        steps += 1
        el = time.perf_counter() - t0
        if el * (steps + 1) / steps > budget_s or steps >= 3:
            break
    # the reference as shipped runs one rank without threads: one step on one core
    O.set_threads(1)
    t0 = time.perf_counter()
    O.advance(F, psi0, geo, stepper="RK4", dt=5e-4, imaginary=False, nsteps=1)
    one = time.perf_counter() - t0
    O.set_threads(cores)
    return {
        "value": steps / el,
        "unit": "steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{steps} RK4 step(s) of the {shape[0]}^3 droplet workload, CPU oracle "
                  f"(C restatement of the reference FFTW path) with OpenMP over {cores} threads",
        "one_thread": {"value": 1.0 / one, "unit": "steps/s", "cores": 1, "sample": "1 RK4 step, 1 thread"},
    }


def cpu_ground_state_subset(n, L, N, psi0, gpu_steps, gpu_en, dt_factor, tol, check_every, budget_s: float = 90.0):
    """A bounded CPU subset of the batched ground states: the 64^3 instance that the GPU
    converged in the fewest steps, run cold by the CPU oracle (OpenMP, up to 16 cores)
    with the same rule (energy at the start of the last step of every check_every
    steps, converged when it moved by at most tol relative since the previous check).
    Reports the CPU time and steps/s, and the GPU's steps and energy for the same
    instance (they must agree: same step count, energy to 1e-10)."""
    from oracle import oracle as O

    O.build()
    cores = min(_affinity(), 16)
    O.set_threads(cores)
    cand = [i for i in range(len(n)) if n[i] == 64 and gpu_steps[i] > 0]
    if not cand:
        return None
    i = min(cand, key=lambda k: gpu_steps[k])
    dx = L[i] / n[i]
    geo = O.Geometry((n[i],) * 3, (-L[i] / 2,) * 3, (L[i] / 2,) * 3, 1)
    F = O.Functional(kind=O.OR_LHY)
    kw = dict(stepper="RK4", dt=dt_factor * dx * dx, imaginary=True, renormalize=[N[i]])
    a = np.asfortranarray(psi0[i][..., None])
    t0 = time.perf_counter()
    steps, prev, E, converged = 0, None, None, False
    while time.perf_counter() - t0 < budget_s:
        a = O.advance(F, a, geo, nsteps=check_every - 1, **kw)
        E = O.energy(F, a, geo)
        a = O.advance(F, a, geo, nsteps=1, **kw)
        steps += check_every
        if prev is not None and abs(E - prev) <= tol * abs(E):
            converged = True
            break
        prev = E
    el = time.perf_counter() - t0
    return {"sample": f"one 64^3 droplet of the batch (instance {i}, N = {N[i]:.1f}) to {tol:g} on the CPU oracle "
                      f"with OpenMP over {cores} threads",
            "cores": cores, "seconds": el, "steps": steps, "converged": converged,
            "steps_per_s": steps / el, "energy": E, "gpu_steps": int(gpu_steps[i]), "gpu_energy": float(gpu_en[i]),
            "energy_rel_diff_vs_gpu": abs(E - gpu_en[i]) / abs(gpu_en[i])}


def ground_state_batch(count: int, seed: int = 3, dt_factor: float = 0.1, tol: float = 1e-10,
                       check_every: int = 100, max_steps: int = 200000, cpu: bool = True):
    """Cold-started batched LHY droplet ground states: `count` instances with N log-uniform
    in [50, 5000] (seeded), grid per instance from {64, 96, 128, 192}^3 by the predicted
    radius (box = 6 radii at dx = 0.35), Gaussian seeds, imaginary-time RK4 with per-stage
    renormalisation (dt = 0.1 dx^2), until each instance's energy changes by at most 1e-10
    relative between checks 100 steps apart.  Timed: gp5_batch_ground_state (wall clock,
    graph capture and host-side convergence checks included); the autotuner runs before."""
    from gp5_amd import capi, synth

    n, L, N, psi0 = synth.droplet_batch(count, seed)
    b = capi.Batch(n, L, N)
    for i, a in enumerate(psi0):
        b.upload(i, a)
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    t0 = time.perf_counter()
    report = b.tune(f, dt_factor, 2)
    tune_s = time.perf_counter() - t0
    _progress(f"cold ground states of {count} droplets")
    steps, en, ms = b.ground_state(f, dt_factor, tol, check_every, max_steps)
    b.close()
    _progress(f"cold: {ms / 1e3:.1f} s; surrogate-seeded run")
    cpu_gs = cpu_ground_state_subset(n, L, N, psi0, steps, en, dt_factor, tol, check_every) if cpu else None
    seeded = surrogate_seeded(n, L, N, f, dt_factor, tol, check_every, max_steps, ms, steps, en)
    _progress("two-level runs")
    two = two_level_legs(n, L, N, f, dt_factor, tol, check_every, max_steps, ms, steps, en)
    conv = steps > 0
    inst_steps = int(np.abs(steps).sum())
    pts = np.array([k ** 3 for k in n], dtype=np.float64)
    out = {
        "workload": f"{count} LHY droplets, N log-uniform [50, 5000] (seed {seed}), grids "
                    + ", ".join(f"{g}^3 x{int(sum(1 for k in n if k == g))}" for g in synth.BATCH_GRIDS)
                    + f", dx = {synth.BATCH_DX}, imaginary-time RK4 dt = {dt_factor} dx^2 with per-stage "
                      f"renormalisation, converged to {tol:g} relative energy change per {check_every} steps",
        "seconds": ms / 1e3,
        "converged": int(conv.sum()),
        "instances": count,
        "steps_min": int(np.abs(steps).min()), "steps_median": float(np.median(np.abs(steps))),
        "steps_max": int(np.abs(steps).max()),
        "instance_steps": inst_steps,
        "point_steps_per_s": float((np.abs(steps) * pts).sum() / (ms / 1e3)),
        "energy_per_atom_min": float((en / np.array(N)).min()),
        "energy_per_atom_max": float((en / np.array(N)).max()),
        "tune_seconds": tune_s,
        "tuning": report.strip().splitlines(),
        "start": "cold (Gaussian seeds)",
        "cpu_subset": cpu_gs,
        "surrogate_seeded": seeded,
        "two_level": two,
    }
    # speed-up of every start over the cold run, same convergence rule and energies
    out["speedup_vs_cold"] = {
        "network_seed_only": seeded["speedup_vs_cold"] if seeded else None,
        "two_level_gaussian_coarse": two.get("gaussian", {}).get("speedup_vs_cold"),
        "two_level_network_coarse": two.get("surrogate", {}).get("speedup_vs_cold"),
    }
    return out


def surrogate_seeded(n, L, N, f, dt_factor, tol, check_every, max_steps, cold_ms, cold_steps, cold_en):
    """The same batch seeded by the trained network (gp5_amd/data/surrogate.npz): seeds are
    predicted on the GPU (bf16 MFMA U-Net) and mapped to each grid on the host; the
    ground-state run is timed like the cold one."""
    from gp5_amd import capi, surrogate

    if not os.path.exists(surrogate.DEFAULT_WEIGHTS):
        return None
    t0 = time.perf_counter()
    sur = surrogate.Surrogate()
    seeds = [sur.seed(N[i], (0.0, 0.0, 0.0), 2.5, n[i], L[i]) for i in range(len(n))]
    seed_s = time.perf_counter() - t0
    b = capi.Batch(n, L, N)
    for i, a in enumerate(seeds):
        b.upload(i, a)
    b.tune(f, dt_factor, 2)
    steps, en, ms = b.ground_state(f, dt_factor, tol, check_every, max_steps)
    b.close()
    return {
        "seconds": ms / 1e3,
        "seeding_seconds": seed_s,
        "converged": int((steps > 0).sum()),
        "instance_steps": int(np.abs(steps).sum()),
        "steps_median": float(np.median(np.abs(steps))),
        "speedup_vs_cold": cold_ms / ms,
        "speedup_vs_cold_incl_seeding": (cold_ms / 1e3) / (ms / 1e3 + seed_s),
        "instance_steps_ratio": float(np.abs(cold_steps).sum() / np.abs(steps).sum()),
        "max_energy_per_atom_diff_vs_cold": float(np.max(np.abs(en - cold_en) / np.array(N))),
    }


def two_level_legs(n, L, N, f, dt_factor, tol, check_every, max_steps, cold_ms, cold_steps, cold_en):
    """The same batch through gp5_amd.ground_state.two_level: ground states on the
    half-resolution grid of each box, resampled on the GPU to the full grid and
    finished there to the same tolerance.  Coarse starts: Gaussians, then the
    learned warm start evaluated on the coarse grid.  Timed: seeding + coarse run +
    resampling + fine run (wall clock); tuning apart, as for the cold run."""
    from gp5_amd import ground_state, surrogate

    out = {}
    legs = [("gaussian", {})]
    warm_s = None
    if os.path.exists(surrogate.DEFAULT_WEIGHTS):
        sur = surrogate.Surrogate(batch=16)  # 16 droplets per forward pass
        # one forward pass before the timed legs: loads the network's kernels (a
        # one-time cost of the process, reported apart like the tuning)
        t0 = time.perf_counter()
        sur.predict_many([(100.0, (0.0, 0.0, 0.0), 2.5)])
        warm_s = time.perf_counter() - t0
        legs.append(("surrogate", {"seed_many": lambda items: sur.seed_many(
            [(Nk, (0.0, 0.0, 0.0), 2.5, k, Lk) for _, k, Lk, Nk in items])}))
    for name, kw in legs:
        fb, rep = ground_state.two_level(n, L, N, f, dt_factor=dt_factor, tol=tol, check_every=check_every,
                                         max_steps=max_steps, **kw)
        fb.close()
        steps = rep["steps"]
        out[name] = {
            "seconds": rep["total_s"],
            "seeding_seconds": rep["seed_s"], "coarse_seconds": rep["coarse_s"],
            "resample_seconds": rep["resample_s"], "fine_seconds": rep["fine_s"], "tune_seconds": rep["tune_s"],
            "coarse_grids": sorted(set(rep["coarse_grids"])),
            "converged": int((steps > 0).sum()),
            "coarse_steps_median": float(np.median(np.abs(rep["coarse_steps"]))),
            "fine_steps_median": float(np.median(np.abs(steps))),
            "fine_instance_steps": int(np.abs(steps).sum()),
            "speedup_vs_cold": (cold_ms / 1e3) / rep["total_s"],
            "max_energy_per_atom_diff_vs_cold": float(np.max(np.abs(rep["energy"] - cold_en) / np.array(N))),
        }
        if name == "surrogate":
            out[name]["network_warmup_seconds"] = warm_s
        _progress(f"two-level ({name}): {rep['total_s']:.1f} s")
    return out


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix peak


def network_mfma(batch: int = 8, steps: int = 10, D: int = 64):
    """MFMA throughput of the warm-start network (3D U-Net, bf16 MFMA with float
    accumulation) at D^3 with `batch` samples: forward passes and training steps
    (forward, backward, Adam) on synthetic inputs.  Model flops: the sum over the
    eight 3x3x3 convolutions of 2 * 27 * Cin * Cout * D_layer^3 * batch for a forward
    pass; a training step adds the weight gradients of all eight and the data
    gradients of c2..c8 (the input layer's data gradient is never computed), so three
    times the forward count less c1's share; pooling, upsampling and the elementwise
    kernels count 0.  The roofline fractions count the MFMA flops the kernels
    execute: the forward of a decoder layer (c5, c6, c7: 32 or 64 input channels at
    side 32..128, 128 at 16..64) runs by output parity, 16 (coarse tap, x parity) pairs of hi and
    lo MFMAs for two fine positions, and the pooled data gradients of c6 and c7 run as
    stride-2 convolutions, 64 folded taps of hi and lo MFMAs for the 8 children of a
    cell: 32/54 of the model count for each; the output layer's weight gradient, whose
    dY has one nonzero channel, runs the taps as the MFMA's N dimension and its data
    gradient the taps as K (1024 of the model's 13824 flops per position each); the decoder layers' weight gradients sum 64
    products per coarse cell for the 27 taps of its 8 children (8/27).  Wall time between device
    synchronisations (the loss stays on the device), so launch costs are included."""
    from gp5_amd import nn

    div = {"c1": 1, "c2": 2, "c3": 4, "c4": 8, "c5": 4, "c6": 2, "c7": 1, "c8": 1}
    layer = {name: 2.0 * 27 * ci * co * (D // div[name]) ** 3 * batch for name, ci, co in nn.SHAPES}
    model_fwd = sum(layer.values())
    model_train = 3 * model_fwd - layer["c1"]
    parity = os.environ.get("GP5_CONV_UP2P", "1") != "0"
    by_parity = {name for name, ci, co in nn.SHAPES
                 if parity and name in ("c5", "c6", "c7")
                 and ((ci in (32, 64) and D // div[name] in (32, 64, 128)) or (ci == 128 and D // div[name] in (16, 32, 64)))}
    stride2 = {name for name, ci, co in nn.SHAPES
               if name in ("c5", "c6", "c7") and nn._pool_fused(D // div[name], ci, co, batch)
               and nn._pool2p(D // div[name], co)}
    fwd_flops = sum(f * (32.0 / 54.0 if name in by_parity else 1.0) for name, f in layer.items())
    dgrad_flops = sum(f * (32.0 / 54.0 if name in stride2 else 1.0) for name, f in layer.items() if name != "c1")
    # the output layer's data gradient from channel 0 alone (gp5_nn_conv3d_act_bwd_out0):
    # one 16x16x32 MFMA per 16 positions, 1024 flops per position against 13824
    if nn._out0c(D):
        dgrad_flops -= layer["c8"] * (1.0 - 1024.0 / 13824.0)
    # the output layer's weight gradient from channel 0 (gp5_nn_conv3d_wgrad_out0): one
    # 16x16x32 MFMA per 16 taps and 32 positions, 1024 flops per position against 13824
    out0 = D in (32, 64, 128) and os.environ.get("GP5_WGRAD_OUT0", "1") != "0"
    # the decoder layers' weight gradients by output parity (k_wgrad_up2): 64 products
    # G_a[o] per coarse cell against 216 = 27 taps x 8 children, 8/27 of the model count
    wup2 = {name for name in ("c5", "c6", "c7")
            if D // div[name] in (16, 32, 64, 128) and os.environ.get("GP5_WGRAD_UP2", "1") != "0"}
    wgrad_flops = (sum(f * (8.0 / 27.0 if name in wup2 else 1.0) for name, f in layer.items())
                   - (layer["c8"] * (1.0 - 1024.0 / 13824.0) if out0 else 0.0))
    train_flops = fwd_flops + dgrad_flops + wgrad_flops
    net = nn.UNet(batch, D, seed=0)
    rng = np.random.default_rng(0)
    X = (0.5 * rng.standard_normal((batch, D, D, D, nn.C0))).astype(np.float32)
    base = np.ascontiguousarray(X[..., 0])
    target = (base + 0.01 * rng.standard_normal(base.shape)).astype(np.float32)
    net.set_batch(X, base, target)
    net.forward()
    net.backward()
    net.step(1e-4)
    net.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward()
    net.sync()
    fwd_s = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        net.forward()
        net.backward(read_loss=False)  # the loss stays on the device (summed there)
        net.step(1e-4)
    net.sync()
    train_s = (time.perf_counter() - t0) / steps
    fwd_tf = fwd_flops / fwd_s / 1e12
    train_tf = train_flops / train_s / 1e12
    return {
        "workload": f"3D U-Net 16-32-64-128 channels, {D}^3, batch {batch}, bf16 MFMA, float accumulation",
        "forward_flops": fwd_flops, "forward_ms": 1e3 * fwd_s, "forward_TFLOPs": fwd_tf,
        "train_step_flops": train_flops, "train_step_ms": 1e3 * train_s, "train_TFLOPs": train_tf,
        "model_forward_flops": model_fwd, "model_forward_TFLOPs": model_fwd / fwd_s / 1e12,
        "model_train_step_flops": model_train, "model_train_TFLOPs": model_train / train_s / 1e12,
        "by_output_parity": sorted(by_parity), "data_gradient_stride2": sorted(stride2),
        "weight_gradient_by_parity": sorted(wup2),
        "roofline": {"bound": "mfma", "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "forward_frac": fwd_tf / MFMA_BF16_PEAK_TFLOPS, "train_frac": train_tf / MFMA_BF16_PEAK_TFLOPS,
                     "flops": "executed MFMA flops",
                     # the same times against the direct 3x3x3 convolution's flop count
                     "model_forward_frac": model_fwd / fwd_s / 1e12 / MFMA_BF16_PEAK_TFLOPS,
                     "model_train_frac": model_train / train_s / 1e12 / MFMA_BF16_PEAK_TFLOPS},
    }


def collision_large(steps: int, n: int = 512):
    """Two N = 4000 LHY droplets colliding head-on in a 512^3 periodic box L = 64
    (dx = 0.125), centres x = -8 and +8, momenta +-0.49 (0.5 snapped to 2 pi m / L);
    real-time RK4, dt = 2.5e-4.  The droplets are one ground state of the box
    (learned warm start on the 256^3 coarse grid, FP64 imaginary time to 1e-10 on
    both levels, gp5_amd.ground_state.droplet_pair), moved to +-8.  Timed like the
    headline run: `steps` steps between stream synchronisations, after 2 warmup
    steps; norm and energy from the fused reductions before and after."""
    from gp5_amd import capi, ground_state

    _progress(f"two-droplet collision at {n}^3: ground state of one droplet")
    shape, left, right, psi0, q, init = ground_state.droplet_pair(n)
    _progress(f"ground state in {init['seconds']:.1f} s; collision")
    ctx = capi.Context(shape, left, right, 1)
    psi = ctx.field(psi0)
    del psi0
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    s = capi.stepper_desc(capi.GP5_STEP_RK4, dt=2.5e-4, imaginary=False)
    norms0, ens0 = capi.advance(ctx, f, s, psi, 1, log=True)
    capi.advance(ctx, f, s, psi, 1)
    ctx.sync()
    t0 = time.perf_counter()
    dev_ms = capi.advance_timed(ctx, f, s, psi, steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    norms1, ens1 = capi.advance(ctx, f, s, psi, 1, log=True)
    F = 16.0 * n ** 3
    sps = steps / wall
    out = {
        "workload": f"two LHY droplets N=4000 each colliding, {n}^3 periodic box L=64, x0=+-8, "
                    f"k=+-{q:.4f}, real-time RK4 dt=2.5e-4, complex128",
        "steps": steps, "steps_per_s": sps, "ms_per_step": 1e3 * wall / steps,
        "device_ms_per_step": dev_ms / steps,
        "model_bytes_per_step": STEP_MODEL_F * F,
        "frac_of_hbm_peak": STEP_MODEL_F * F * sps / 1e9 / HBM_PEAK_GBPS,
        "norm_start": float(norms0[0, 0]), "norm_end": float(norms1[0, 0]),
        "energy_start": float(ens0[0]), "energy_end": float(ens1[0]),
        "steps_between": steps + 2,
        "initial_state": init,
    }
    psi.free()
    ctx.close()
    return out


def config1_harmonic(gpu_steps: int = 2000, cpu: bool = True, cpu_budget_s: float = 20.0):
    """The reference's CPU-runnable case: single-component GP (g = 1, mass 1), 64^3 on
    [-10, 10]^3, harmonic trap omega = (1, 1, 1), N = 1, Gaussian sigma = 1.5 plus
    seeded 1e-3 complex noise, imaginary-time Euler dt = 1e-3 renormalised every step,
    2000 steps.  GPU steps/s over the 2000 steps (field resident in HBM); GPU vs the
    CPU oracle after 100 steps (relative L2); the oracle's steps/s on all host cores
    (up to 16, the 2000 steps in blocks of 100 within cpu_budget_s) and on one core
    (the reference as shipped: one rank, FFTW without threads; 100 steps); mu and E
    of the final state."""
    from gp5_amd import capi, synth

    shape, left, right, psi0, V = synth.harmonic_64()
    ctx = capi.Context(shape, left, right, 1)
    vb = ctx.real(V)
    f = capi.functional_desc(capi.GP5_FUNC_GP, couplings=[[1.0]], potential=vb)
    s = capi.stepper_desc(capi.GP5_STEP_EULERO, dt=1e-3, imaginary=True, renormalize=[1.0])
    psi = ctx.field(psi0)
    capi.advance(ctx, f, s, psi, 100)
    got100 = psi.download_complex()
    psi.upload_complex(psi0)
    capi.advance(ctx, f, s, psi, 20)  # warm-up
    psi.upload_complex(psi0)
    ctx.sync()
    t0 = time.perf_counter()
    capi.advance(ctx, f, s, psi, gpu_steps)
    ctx.sync()
    gpu_s = time.perf_counter() - t0
    # mu = <psi|H|psi> / N of the final state, E from the energy functional
    H = ctx.field()
    capi.functional_apply(ctx, f, psi, H)
    a, h = psi.download_complex(), H.download_complex()
    dV = synth.dV(shape, left, right)
    out = {"workload": "GP g=1, 64^3 on [-10,10]^3, harmonic omega=(1,1,1), N=1, Gaussian sigma=1.5 + 1e-3 noise "
                       "(seed 1), imaginary-time Euler dt=1e-3, renormalised every step, 2000 steps",
           "gpu_steps_per_s": gpu_steps / gpu_s, "gpu_steps": gpu_steps,
           "mu": float(np.real(np.vdot(a, h)) * dV), "energy": capi.energy(ctx, f, psi)}
    if cpu:
        from oracle import oracle as O

        O.build()
        cores = min(_affinity(), 16)
        O.set_threads(cores)
        geo = O.Geometry(shape, left, right, 1)
        F = O.Functional(kind=O.OR_GP, couplings=np.array([[1.0]]), V=V)
        t0 = time.perf_counter()
        ref = O.advance(F, psi0, geo, stepper="eulero", dt=1e-3, imaginary=True, nsteps=100, renormalize=[1.0])
        done, el = 100, time.perf_counter() - t0
        out["rel_l2_vs_cpu_at_step_100"] = float(np.linalg.norm((got100 - ref).ravel()) / np.linalg.norm(ref.ravel()))
        cur = ref
        while done < gpu_steps and el * (done + 100) / done < cpu_budget_s:
            cur = O.advance(F, cur, geo, stepper="eulero", dt=1e-3, imaginary=True, nsteps=100, renormalize=[1.0])
            done += 100
            el = time.perf_counter() - t0
        out["cpu_all_cores"] = {"value": done / el, "unit": "steps/s", "cores": cores, "steps": done,
                                "sample": f"{done} of the 2000 Euler steps, CPU oracle with OpenMP over {cores} threads"}
        if done == gpu_steps:
            out["rel_l2_vs_cpu_at_step_2000"] = float(np.linalg.norm((a - cur).ravel()) / np.linalg.norm(cur.ravel()))
        O.set_threads(1)
        t0 = time.perf_counter()
        O.advance(F, psi0, geo, stepper="eulero", dt=1e-3, imaginary=True, nsteps=100, renormalize=[1.0])
        one = time.perf_counter() - t0
        O.set_threads(cores)
        out["cpu_one_core"] = {"value": 100 / one, "unit": "steps/s", "cores": 1, "steps": 100,
                               "sample": "100 Euler steps, CPU oracle on one thread"}
    H.free()
    psi.free()
    vb.free()
    ctx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--gs-count", type=int, default=64,
                    help="instances of the batched ground-state workload (0 skips it)")
    ap.add_argument("--nn-batch", type=int, default=8,
                    help="batch of the warm-start network's MFMA measurement (0 skips it)")
    ap.add_argument("--large-steps", type=int, default=400,
                    help="RK4 steps of the 512^3 two-droplet collision (0 skips it)")
    ap.add_argument("--no-config1", dest="config1", action="store_false",
                    help="skip the 64^3 harmonic-trap Euler case (GPU and CPU steps/s)")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with one normalize of the field: its scale kernel reads and writes the field "
                         "once, the reference for rocprofv3 FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    from gp5_amd import capi, synth

    n = args.n
    shape, left, right, psi0 = synth.droplet_256(n)
    ctx = capi.Context(shape, left, right, 1)
    psi = ctx.field(psi0)
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    s = capi.stepper_desc(capi.GP5_STEP_RK4, dt=5e-4, imaginary=False)

    _progress(f"RK4 at {n}^3")
    # norm / energy at the start (fused reductions of stage 1)
    norms0, ens0 = capi.advance(ctx, f, s, psi, 1, log=True)
    if args.warmup > 1:
        capi.advance(ctx, f, s, psi, args.warmup - 1)

    ctx.sync()
    t0 = time.perf_counter()
    dev_ms = capi.advance_timed(ctx, f, s, psi, args.steps)
    ctx.sync()
    wall = time.perf_counter() - t0

    steps_per_s = args.steps / wall
    ms_per_step = 1e3 * wall / args.steps

    # per-kernel timing (HIP events around each launch) for the roofline object
    prof = capi.profile_advance(ctx, f, s, psi, args.profile_steps)
    # N(t) and E(t) of the same workload from its start, every 10 steps of args.steps
    # (fused stage-1 reductions; outside the timed region)
    logf = ctx.field(psi0)
    nlog, elog = capi.advance(ctx, f, s, logf, args.steps, log=True)
    logf.free()
    norms1, ens1 = capi.advance(ctx, f, s, psi, 1, log=True)
    F = 16.0 * shape[0] * shape[1] * shape[2]
    dom = max((k for k in prof if k in KIND_BYTES_F), key=lambda k: prof[k][0])
    avg_ms = prof[dom][0] / prof[dom][1]
    bytes_launch = KIND_BYTES_F[dom] * F
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(dom)
        except Exception:
            traffic = None

    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"LHY droplet real-time RK4, {n}^3 periodic box L=40, complex128, "
                               f"flat top r=6 N=3000, dt=5e-4", "grid": n, "components": 1,
                   "stepper": "RK4", "time": "real", "dt": 5e-4},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": bytes_launch},
        "step_roofline": {"model_bytes_per_step": STEP_MODEL_F * F,
                          "achieved_GBps": STEP_MODEL_F * F * steps_per_s / 1e9,
                          "frac": STEP_MODEL_F * F * steps_per_s / 1e9 / HBM_PEAK_GBPS},
        "device_ms_per_step": dev_ms / args.steps,
        "kernels": {k: {"avg_ms": v[0] / v[1], "launches": v[1]} for k, v in prof.items()},
        "checks": {"norm_start": float(norms0[0, 0]), "norm_end": float(norms1[0, 0]),
                   "energy_start": float(ens0[0]), "energy_end": float(ens1[0]),
                   "steps_between": args.warmup + args.steps + args.profile_steps},
        "conservation": {"every_steps": 10, "steps": args.steps,
                         "norm": [float(v) for v in nlog[::10, 0]], "energy": [float(v) for v in elog[::10]],
                         "norm_rel_drift_max": float(np.max(np.abs(nlog[:, 0] - nlog[0, 0])) / nlog[0, 0]),
                         "energy_rel_drift_max": float(np.max(np.abs(elog - elog[0])) / abs(elog[0]))},
    }
    # host <-> HBM copies of the field through the boundary (gp5_buf_upload / download of
    # a pageable numpy array); outside the timed region, reported for the PCIe-inclusive rate
    t0 = time.perf_counter()
    host = psi.download_complex()
    t_d2h = time.perf_counter() - t0
    t0 = time.perf_counter()
    psi.upload_complex(host)
    ctx.sync()
    t_h2d = time.perf_counter() - t0
    out["host_transfer"] = {"field_bytes": F, "d2h_ms": 1e3 * t_d2h, "h2d_ms": 1e3 * t_h2d,
                            "steps_per_s_incl_one_roundtrip": args.steps / (wall + t_d2h + t_h2d)}
    if args.calibrate:
        capi.normalize(ctx, psi, 0, 3000.0)
        ctx.sync()
    if args.config1:
        _progress("config 1: 64^3 harmonic trap, Euler imaginary time")
        out["config1"] = config1_harmonic(cpu=not args.no_cpu)
    if args.gs_count > 0:
        out["ground_state"] = ground_state_batch(args.gs_count, cpu=not args.no_cpu)
    if args.large_steps > 0:
        out["large"] = collision_large(args.large_steps)
    if args.nn_batch > 0:
        _progress("warm-start network MFMA throughput")
        out["network"] = network_mfma(args.nn_batch)
    if not args.no_cpu:
        _progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(psi0, shape, left, right)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
