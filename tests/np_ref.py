"""Independent numpy restatement of GP5's pseudo-spectral step, used to
cross-check the C oracle (it shares no code with it: numpy's pocketfft for the
transforms, numpy.fft.fftfreq for the wavenumbers, vectorised algebra).

Follows src/operators.cpp:7-36, src/functional.cpp:13-150,
src/stepper.cpp:7-81, src/geometry.cpp:190-237, src/constraint.cpp:6-9.
"""
from __future__ import annotations

import numpy as np


def kgrid(shape, left, right):
    """K^2 with k_d = 2 pi fftfreq(N_d, L_d/N_d), i.e. the rule of geometry.cpp:127-148."""
    ks = [2 * np.pi * np.fft.fftfreq(shape[d], d=(right[d] - left[d]) / shape[d]) for d in range(3)]
    return ks[0][:, None, None] ** 2 + ks[1][None, :, None] ** 2 + ks[2][None, None, :] ** 2


def laplacian(psi, shape, left, right):
    K2 = kgrid(shape, left, right)
    out = np.empty_like(psi)
    for c in range(psi.shape[3]):
        out[..., c] = np.fft.ifftn(-K2 * np.fft.fftn(psi[..., c]))
    return out


def apply(psi, shape, left, right, kind="lhy", couplings=None, masses=(1.0, 1.0), a=-3.0, b=2.5,
          alpha=0.0, beta=0.0, eta=0.0, V=None):
    lap = laplacian(psi, shape, left, right)
    rho = np.abs(psi) ** 2
    nc = psi.shape[3]
    out = np.empty_like(psi)
    if kind == "gp":
        for c in range(nc):
            out[..., c] = -0.5 / masses[c] * lap[..., c]
        if V is not None:
            out += V * psi
        if couplings is not None:
            g = np.asarray(couplings, dtype=float)
            if nc == 1:
                out += g[0, 0] * rho * psi
            else:
                out[..., 0] += (g[0, 0] * rho[..., 0] + g[0, 1] * rho[..., 1]) * psi[..., 0]
                out[..., 1] += (g[1, 1] * rho[..., 1] + g[0, 1] * rho[..., 0]) * psi[..., 1]
    elif kind == "lhy":
        out = -0.5 * lap + (a * rho + b * rho ** 1.5) * psi
        if V is not None:
            out += V * psi
    elif kind == "unlocked":
        r1, r2 = rho[..., 0], rho[..., 1]
        s = (r1 + beta * r2) ** 1.5
        out[..., 0] = -0.5 * lap[..., 0] + (r1 + eta * r2 + alpha * s) * psi[..., 0]
        out[..., 1] = -0.5 * lap[..., 1] + (eta * beta * r1 + beta * r2 + alpha * beta ** 2 * s) * psi[..., 1]
        if V is not None:
            out += V * psi
    else:
        raise ValueError(kind)
    return out


def norm(psi, shape, left, right, c=0):
    dV = np.prod([(right[d] - left[d]) / shape[d] for d in range(3)])
    return float(np.sum(np.abs(psi[..., c]) ** 2) * dV)


def advance(psi, shape, left, right, *, stepper="RK4", dt=1e-3, imaginary=True, nsteps=1, renormalize=None,
            phase=None, **fkw):
    tau = dt if imaginary else 1j * dt

    def H(f):
        return apply(f, shape, left, right, **fkw)

    def renorm(f):
        if renormalize is None:
            return f
        f = f.copy()
        for c, N in enumerate(renormalize):
            f[..., c] *= np.sqrt(N / norm(f, shape, left, right, c))
        return f

    def constrain(f):
        if phase is None:
            return f
        return np.abs(f) * np.exp(1j * phase)

    psi = np.array(psi, dtype=np.complex128)
    for _ in range(nsteps):
        if stepper == "eulero":
            psi = renorm(constrain(psi - tau * H(psi)))
        else:
            k1 = H(psi)
            p = renorm(psi - tau * 0.5 * k1)
            k2 = H(p)
            p = renorm(psi - tau * 0.5 * k2)
            k3 = H(p)
            p = renorm(psi - tau * k3)
            k4 = H(p)
            psi = renorm(constrain(psi - (tau / 6) * (((k1 + 2 * k2) + 2 * k3) + k4)))
    return psi
