"""YAML-configured run: the role of the reference's `gp` executable
(src/main.cpp:14-249) on one MI355X.

    python -m gp5_amd.driver input.yaml

Reads the same input files (exampleInputFiles/*.yaml): mesh, domain, parallel,
wavefunction, functional (gpFunctional / LHYDroplet / UnlockedLHYDroplet /
externalPotential), evolution (timeStep, imaginaryTime, stepper, maxTime,
reNormalize, constraint) and initialCondition / output.  The initial state is
an HDF5 field file; out_<k>.hdf5 is written every nIterations steps while
t <= maxTime, with the reference's loop and its "Time: t" / "saved" lines on
stdout.  Every step runs in the HIP kernels of libgp5.so (gp5_advance); the
state stays in HBM between output blocks.

Differences from the reference, all in what it would reject or crash on:
  * parallel.processorGrid must multiply to 1 (one GPU); fftOrdering is
    optional and ignored (it only permutes the distributed k-space layout);
  * evolution.reNormalize is optional (default false); the reference requires
    it, and its own exampleInputFiles/dropletUnlocked.yaml omits it;
  * an unknown stepper name, a functional section with only potentials, a
    complex potential or phase file, or more than two components raise
    ValueError instead of crashing.
"""
from __future__ import annotations

import os
import sys
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import yaml

from . import capi


@dataclass
class RunConfig:
    shape: List[int]
    left: List[float]
    right: List[float]
    nc: int
    normalizations: List[float]
    functional: dict
    time_step: float
    imaginary: bool
    stepper: str
    max_time: float
    renormalize: bool
    phase_file: Optional[str]
    initial_file: str
    out_dir: str
    n_iterations: int
    raw: dict = field(default_factory=dict, repr=False)


def _req(node, key, where):
    if not isinstance(node, dict) or key not in node:
        raise ValueError(f"missing '{key}' in {where}")
    return node[key]


def parse(cfg: dict) -> RunConfig:
    """Read the sections main.cpp reads (src/main.cpp:32-166)."""
    shape = [int(v) for v in _req(_req(cfg, "mesh", "input"), "shape", "mesh")]
    dom = _req(cfg, "domain", "input")
    left = [float(v) for v in _req(dom, "left", "domain")]
    right = [float(v) for v in _req(dom, "right", "domain")]
    if len(shape) != 3 or len(left) != 3 or len(right) != 3:
        raise ValueError("mesh.shape, domain.left and domain.right need 3 entries")
    wf = _req(cfg, "wavefunction", "input")
    nc = int(_req(wf, "nComponents", "wavefunction"))
    if nc not in (1, 2):
        raise ValueError("nComponents must be 1 or 2")
    norms = [float(v) for v in _req(wf, "normalization", "wavefunction")]
    par = cfg.get("parallel") or {}
    grid = [int(v) for v in par.get("processorGrid", [1, 1, 1])]
    if int(np.prod(grid)) != 1:
        raise ValueError(f"processorGrid {grid}: this build runs one GPU, use [1, 1, 1]")
    ev = _req(cfg, "evolution", "input")
    constraint = ev.get("constraint")
    phase_file = None
    if constraint:
        # constraintConstructor::create (src/constraint.cpp:19-38): the first
        # phaseConstraint entry wins, other names are ignored
        for entry in constraint.values():
            if entry.get("name") == "phaseConstraint":
                phase_file = str(_req(entry, "file", "phaseConstraint"))
                break
    out = _req(cfg, "output", "input")
    return RunConfig(
        shape=shape, left=left, right=right, nc=nc, normalizations=norms,
        functional=_req(cfg, "functional", "input"),
        time_step=float(_req(ev, "timeStep", "evolution")),
        imaginary=bool(_req(ev, "imaginaryTime", "evolution")),
        stepper=str(_req(ev, "stepper", "evolution")),
        max_time=float(_req(ev, "maxTime", "evolution")),
        renormalize=bool(ev.get("reNormalize", False)),
        phase_file=phase_file,
        initial_file=str(_req(_req(cfg, "initialCondition", "input"), "file", "initialCondition")),
        out_dir=str(_req(out, "folder", "output")),
        n_iterations=int(_req(out, "nIterations", "output")),
        raw=cfg,
    )


def positions(shape: Sequence[int], left: Sequence[float], right: Sequence[float], nc: int):
    """Cell-centred X, Y, Z of shape (Nx, Ny, Nz, nC) (src/geometry.cpp:165-188)."""
    axes = []
    for d in range(3):
        delta = (right[d] - left[d]) / shape[d]
        axes.append(left[d] + delta * (np.arange(shape[d]) + 0.5))
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    rep = lambda A: np.repeat(A[..., None], nc, axis=3)
    return rep(X), rep(Y), rep(Z)


def _real_part(z: np.ndarray, what: str) -> np.ndarray:
    if np.any(np.imag(z) != 0.0):
        raise ValueError(f"{what}: nonzero imaginary part; this build takes real potentials and phases")
    return np.asfortranarray(np.real(z))


def potential(entries: dict, shape, left, right, nc: int) -> np.ndarray:
    """externalPotentialConstructor::create + sumPotential::create
    (src/externalPotential.cpp:58-74, 86-124): the sum of harmonic, vortex and
    potentialFromFile terms in input order, as a real (Nx, Ny, Nz, nC) array."""
    X, Y, Z = positions(shape, left, right, nc)
    V = np.zeros(tuple(shape) + (nc,), order="F")
    for key, p in entries.items():
        name = _req(p, "name", f"potential {key}")
        if name == "harmonic":
            # harmonicPotential::create (src/externalPotential.cpp:11-32)
            om = [[float(w) for w in row] for row in _req(p, "omegas", "harmonic")]
            if len(om) < nc:
                raise ValueError("harmonic: need one omegas triple per component")
            term = np.zeros_like(V)
            for c in range(nc):
                w = om[c]
                x, y, z = X[..., c], Y[..., c], Z[..., c]
                term[..., c] = (w[0] * w[0] * x * x + w[1] * w[1] * y * y + w[2] * w[2] * z * z) * 0.5
        elif name == "vortex":
            # vortexPotential::create (src/externalPotential.cpp:34-55)
            term = 0.5 / (X * X + Y * Y)
        elif name == "potentialFromFile":
            # potentialFromFile::create (src/externalPotential.cpp:76-82)
            fname = str(_req(p, "filename", "potentialFromFile"))
            term = _real_part(capi.h5_read(fname, shape, nc), fname)
        else:
            raise ValueError("Unkown potential")
        V = V + term
    return np.asfortranarray(V)


def functional(cfg_f: dict, shape, left, right, nc: int):
    """functionalConstructor::create (src/functional.cpp:153-258): returns the
    functional descriptor and the real potential array (or None).  As in the
    reference, the last functional entry and the last externalPotential entry
    win."""
    desc = None
    V = None
    for key, fc in cfg_f.items():
        name = _req(fc, "name", f"functional {key}")
        if name == "gpFunctional":
            couplings = None
            masses = [1.0] * nc
            if "coupling" in fc:
                g = [[float(v) for v in row] for row in fc["coupling"]]
                if len(g) != nc or any(len(row) != nc for row in g):
                    raise ValueError("Incompatible number of components")
                # couplings(i, j) = coupling[j][i] (src/functional.cpp:168-194)
                couplings = np.asarray(g, dtype=float).T
            if "masses" in fc:
                # setMasses, then init() resizes to nComponents with 1 (src/functional.cpp:92-102)
                masses = ([float(v) for v in fc["masses"]] + [1.0] * nc)[:nc]
            desc = dict(kind=capi.GP5_FUNC_GP, couplings=couplings, masses=tuple(masses))
        elif name == "LHYDroplet":
            desc = dict(kind=capi.GP5_FUNC_LHY)
        elif name == "UnlockedLHYDroplet":
            if nc != 2:
                raise ValueError("Incompatible number of components")
            desc = dict(kind=capi.GP5_FUNC_LHY_UNLOCKED, alpha=float(_req(fc, "alpha", key)),
                        beta=float(_req(fc, "beta", key)), eta=float(_req(fc, "eta", key)))
        elif name == "externalPotential":
            pass
        else:
            raise ValueError(" Unkown functional")
    if desc is None:
        raise ValueError("the functional section needs gpFunctional, LHYDroplet or UnlockedLHYDroplet")
    for key, fc in cfg_f.items():
        if fc["name"] == "externalPotential":
            V = potential(_req(fc, "V", f"functional {key}") or {}, shape, left, right, nc)
    return desc, V


def output_times(time_step: float, n_iterations: int, max_time: float):
    """Times of the outputs of src/main.cpp:180-205: save at t, advance
    nIterations steps adding timeStep to t once per step, repeat while
    t <= maxTime.  The state after the last block is not written."""
    if not time_step > 0 or n_iterations < 1:
        raise ValueError("evolution.timeStep must be > 0 and output.nIterations >= 1")
    t = 0.0
    while t <= max_time:
        yield t
        for _ in range(n_iterations):
            t += time_step


def run(path: str, *, out=None, err=None, max_outputs: Optional[int] = None) -> List[str]:
    """Run an input file; returns the output files written, in order."""
    out = out or sys.stdout
    err = err or sys.stderr
    with open(path) as fh:
        rc = parse(yaml.safe_load(fh))
    t0 = time.perf_counter()
    print("Initializing FFT operator ...", file=err, flush=True)
    ctx = capi.Context(rc.shape, rc.left, rc.right, rc.nc)
    print("Initializing functional ...", file=err, flush=True)
    desc, V = functional(rc.functional, rc.shape, rc.left, rc.right, rc.nc)
    Vbuf = ctx.real(V) if V is not None else None
    f = capi.functional_desc(potential=Vbuf, **desc)
    print("Initializing fields ...", file=err, flush=True)
    psi = ctx.field(capi.h5_read(rc.initial_file, rc.shape, rc.nc))
    print("Initializing stepper ...", file=err, flush=True)
    if rc.stepper == "eulero":
        kind = capi.GP5_STEP_EULERO
    elif rc.stepper == "RK4":
        kind = capi.GP5_STEP_RK4
    else:
        raise ValueError(f"Unknown stepper '{rc.stepper}' (allowed: RK4, eulero)")
    renorm = None
    if rc.renormalize:
        if len(rc.normalizations) < rc.nc:
            raise ValueError("wavefunction.normalization needs one entry per component")
        renorm = rc.normalizations[:rc.nc]
    phase = None
    if rc.phase_file:
        phase = ctx.real(_real_part(capi.h5_read(rc.phase_file, rc.shape, rc.nc), rc.phase_file))
    s = capi.stepper_desc(kind, dt=rc.time_step, imaginary=rc.imaginary, renormalize=renorm, phase=phase)
    os.makedirs(rc.out_dir, exist_ok=True)
    print("START", file=err, flush=True)
    t_io = t_adv = 0.0
    written = []
    for k, t in enumerate(output_times(rc.time_step, rc.n_iterations, rc.max_time)):
        if max_outputs is not None and k >= max_outputs:
            break
        print(f"Time: {t:g}", file=out, flush=True)
        a = time.perf_counter()
        name = os.path.join(rc.out_dir, f"out_{k}.hdf5")
        capi.h5_write(name, psi.download_complex())
        written.append(name)
        t_io += time.perf_counter() - a
        print("saved", file=out, flush=True)
        a = time.perf_counter()
        capi.advance(ctx, f, s, psi, rc.n_iterations)
        t_adv += time.perf_counter() - a
    total = time.perf_counter() - t0
    print(f"timers: total {total:.3f} s, io {t_io:.3f} s, advance {t_adv:.3f} s", file=err, flush=True)
    return written


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print(" The program must have one argument only : the yaml input file", file=sys.stderr)
        return 2
    run(argv[0])
    return 0


if __name__ == "__main__":
    sys.exit(main())
