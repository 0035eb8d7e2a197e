"""Writes the golden fixtures of this directory (inputs and expected outputs).

The reference (lucaparisi91/GP5) cannot be built or imported in this
environment (FFTW3, Eigen Tensor, yaml-cpp, P3DFFT++, MPI and parallel HDF5 are
absent), so the expected outputs come from the CPU oracle (oracle/), itself
pinned by tests/test_oracle.py against the reference's known-answer tests,
closed forms and an independent numpy restatement.  The fixtures freeze those
outputs so the GPU path and later changes of the oracle are checked against the
same numbers.

    python tests/golden/make_golden.py [fixture names]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gp5_amd import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402


def droplet_rk4_real():
    """LHY droplet (-3, 2.5), 16^3 on L = 40, flat top r = 6 N = 3000, real-time RK4 dt = 5e-4, 10 steps."""
    shape, left, right, psi0 = synth.droplet_256(16)
    out = O.advance(O.Functional(kind=O.OR_LHY), psi0, O.Geometry(shape, left, right, 1), stepper="RK4",
                    dt=5e-4, imaginary=False, nsteps=10)
    return dict(shape=shape, left=left, right=right, psi0=psi0, out=out, kind="lhy", stepper="RK4",
                imaginary=False, dt=5e-4, nsteps=10, renorm=np.zeros(0))


def harmonic_euler_imag():
    """gpFunctional g = 1, harmonic trap omega = 1, 16^3 on [-10, 10]^3, Euler imaginary dt = 1e-3,
    renormalised to N = 1, 10 steps."""
    shape, left, right, psi0, V = synth.harmonic_64(16)
    out = O.advance(O.Functional(kind=O.OR_GP, couplings=np.array([[1.0]]), V=V), psi0,
                    O.Geometry(shape, left, right, 1), stepper="eulero", dt=1e-3, imaginary=True, nsteps=10,
                    renormalize=[1.0])
    return dict(shape=shape, left=left, right=right, psi0=psi0, V=V, out=out, kind="gp", couplings=np.array([[1.0]]),
                stepper="eulero", imaginary=True, dt=1e-3, nsteps=10, renorm=np.array([1.0]))


def two_component_rk4_imag():
    """Two-component gpFunctional (g11 1, g12 0.4, g22 0.7, masses 1 and 2) with traps, 12x10x8,
    RK4 imaginary dt = 2e-3 renormalised to (2, 0.5), 8 steps."""
    shape, left, right = (12, 10, 8), (-4.0, -3.5, -3.0), (4.0, 3.5, 3.0)
    r2 = synth.r2_grid(shape, left, right)
    u = synth.uniform(5, 2 * 2 * int(np.prod(shape))).reshape((2, 2) + shape, order="C")
    psi0 = np.empty(shape + (2,), dtype=np.complex128, order="F")
    for c in range(2):
        psi0[..., c] = np.exp(-r2 / (2 * (1.2 + 0.2 * c) ** 2)) * (1 + 0.05 * (u[c, 0] + 1j * u[c, 1]))
    V = synth.harmonic_potential(shape, left, right, [(1.0, 0.9, 1.1), (0.6, 0.7, 0.8)])
    g = np.array([[1.0, 0.4], [0.4, 0.7]])
    out = O.advance(O.Functional(kind=O.OR_GP, couplings=g, masses=(1.0, 2.0), V=V), psi0,
                    O.Geometry(shape, left, right, 2), stepper="RK4", dt=2e-3, imaginary=True, nsteps=8,
                    renormalize=[2.0, 0.5])
    return dict(shape=shape, left=left, right=right, psi0=psi0, V=V, out=out, kind="gp", couplings=g,
                masses=np.array([1.0, 2.0]), stepper="RK4", imaginary=True, dt=2e-3, nsteps=8,
                renorm=np.array([2.0, 0.5]))


def mu_lhy_imaginary():
    """The reference's ground-state regression protocol (python/test/regressionTests/LHY_imaginary.py:8-55):
    LHY droplet N = 370 on 30^3 over [-10, 10]^3, psi0 = exp(-r^2 / 50), RK4 imaginary time with
    dt = 0.01 dx^2; blocks of 1000 (normalize, step) pairs; after each block
    mu = ln(N / norm) / (2 dt), normalize; stop when |mu - mu_previous| <= 1e-9."""
    n, L, N = 30, 10.0, 370.0
    geo = O.Geometry((n, n, n), (-L, -L, -L), (L, L, L), 1)
    X, Y, Z = geo.grid()
    psi = (np.exp(-(X**2 + Y**2 + Z**2) / (2 * 5**2)) + 0j)[..., None]
    psi = O.normalize(psi, geo, 0, N)
    F = O.Functional(kind=O.OR_LHY)
    dt = 0.01 * geo.dx[0] ** 2
    mu, old, history = 1.0e18, 0.0, []
    while abs(mu - old) > 1e-9:
        for _ in range(1000):
            psi = O.normalize(psi, geo, 0, N)
            psi = O.advance(F, psi, geo, stepper="RK4", dt=dt, imaginary=True, nsteps=1)
        old = mu
        mu = 0.5 * np.log(N / O.norm(psi, geo, 0)) / dt
        psi = O.normalize(psi, geo, 0, N)
        history.append(float(mu))
    return dict(mu=float(mu), blocks=len(history), history=history, dt=dt, N=N, shape=[n, n, n],
                left=[-L] * 3, right=[L] * 3, max_density=float(np.max(np.abs(psi)) ** 2))


FIXTURES = {
    "droplet_rk4_real_16": droplet_rk4_real,
    "harmonic_euler_imag_16": harmonic_euler_imag,
    "gp2c_rk4_imag_12x10x8": two_component_rk4_imag,
}


JSON_FIXTURES = {
    "mu_lhy_imaginary_30": mu_lhy_imaginary,
}


def main(names=None):
    O.build()
    O.set_threads(min(8, os.cpu_count() or 1))
    for name, fn in JSON_FIXTURES.items():
        if names and name not in names:
            continue
        d = fn()
        with open(os.path.join(HERE, name + ".json"), "w") as fh:
            json.dump(d, fh, indent=1)
        print(name, d["mu"], d["blocks"])
    for name, fn in FIXTURES.items():
        if names and name not in names:
            continue
        d = fn()
        arrays = {k: np.asarray(v) for k, v in d.items()}
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        print(name, {k: v.shape for k, v in arrays.items() if hasattr(v, "shape")})


if __name__ == "__main__":
    main(sys.argv[1:])
