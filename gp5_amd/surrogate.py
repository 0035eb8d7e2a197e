"""Learned warm start for LHY droplet ground states.

Parameters: atom number N in [50, 5000], trap frequencies w = (wx, wy, wz) in
[0, 0.2], LHY coefficient b in [2, 3] (H psi = -1/2 lap psi + (-3 rho + b rho^{3/2} + V) psi,
V = 1/2 sum w_i^2 x_i^2, capped at V_CAP).  No counterpart in the reference.

Scaled problem.  With the saturation density rho0 = (2.5 / b)^2 (zero pressure of
-3 rho + b rho^{3/2}) and the radius R0 = (3 N / (4 pi rho0))^{1/3}, the field
psi(x) = sqrt(rho0) phi(x / R0) solves, after multiplying H by R0^2, the problem
    -1/2 lap_u phi + R0^2 (-3 rho0 |phi|^2 + b rho0^{3/2} |phi|^3 + V(R0 u)) phi
on the box u in [-3, 3]^3 with norm 4 pi / 3 for every parameter set.  The
ground states of these scaled problems on a 64^3 grid are computed in batches
by the solver (gp5_batch_*: per-instance LHY coefficients, potentials and
norms), and the network (gp5_amd/nn.py) learns phi - phi_base, phi_base being a
flat-top profile of radius 1.  A prediction seeds an instance on any grid by
trilinear interpolation at u = x / R0; imaginary time then refines it.

    python -m gp5_amd.surrogate train --count 2000 --epochs 12 --out gp5_amd/data/surrogate.npz
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import capi, nn

A_LHY = -3.0
U_BOX = 3.0          # scaled half-box
GRID = 64
V_CAP = 5.0          # cap of the trap potential (physical units), far above |mu| ~ 0.5
DT_FACTOR = 0.1      # imaginary-time step / dx^2 (stable for the per-stage renormalised RK4)
SURFACE = 0.7        # width of the baseline flat-top surface (physical units)
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "surrogate.npz")


def rho0(b: float) -> float:
    return (-A_LHY / (1.2 * b)) ** 2


def radius(N: float, b: float) -> float:
    return (3.0 * N / (4.0 * np.pi * rho0(b))) ** (1.0 / 3.0)


def u_axis(n: int = GRID) -> np.ndarray:
    return -U_BOX + (np.arange(n) + 0.5) * (2 * U_BOX / n)


def baseline(N: float, b: float, n: int = GRID) -> np.ndarray:
    u = u_axis(n)
    r = np.sqrt(u[:, None, None] ** 2 + u[None, :, None] ** 2 + u[None, None, :] ** 2)
    R0 = radius(N, b)
    return 0.5 * (1.0 - np.tanh((r - 1.0) * R0 / SURFACE))


def potential_scaled(N: float, w, b: float, n: int = GRID) -> np.ndarray:
    """R0^2 V(R0 u) on the scaled grid (capped physical V)."""
    u = u_axis(n)
    R0 = radius(N, b)
    x2 = [(R0 * u) ** 2 for _ in range(3)]
    V = 0.5 * (w[0] ** 2 * x2[0][:, None, None] + w[1] ** 2 * x2[1][None, :, None] + w[2] ** 2 * x2[2][None, None, :])
    return R0 ** 2 * np.minimum(V, V_CAP)


_GRIDS = {}


def _grids(n: int = GRID):
    if n not in _GRIDS:
        u = u_axis(n).astype(np.float32)
        u2 = u * u
        r = np.sqrt(u2[:, None, None] + u2[None, :, None] + u2[None, None, :])
        _GRIDS[n] = (u2, r)
    return _GRIDS[n]


def features(N: float, w, b: float, n: int = GRID, out: np.ndarray = None) -> np.ndarray:
    """16 input channels on the scaled grid (channels last): the baseline profile, the
    trap shape, |u| and the parameters as constant channels.  With `out` (a float32
    array from an earlier call on the same grid) only the channels that depend on
    (N, w, b) are rewritten."""
    u2, r = _grids(n)
    if out is None:
        X = np.zeros((n, n, n, nn.C0), dtype=np.float32)
        X[..., 7] = r / U_BOX
    else:
        X = out
    R0 = radius(N, b)
    X[..., 0] = 0.5 * (1.0 - np.tanh((r - 1.0) * (R0 / SURFACE)))
    a = [(wi / 0.2) ** 2 / 18.0 for wi in w]
    if any(a) or out is None or X[0, 0, 0, 1] != 0.0:
        X[..., 1] = a[0] * u2[:, None, None] + a[1] * u2[None, :, None] + a[2] * u2[None, None, :]
    X[..., 2] = np.log(N / 50.0) / np.log(100.0)
    X[..., 3] = 2.0 * (b - 2.5)
    X[..., 4] = w[0] / 0.2
    X[..., 5] = w[1] / 0.2
    X[..., 6] = w[2] / 0.2
    return X


def sample_params(count: int, seed: int):
    """N log-uniform [50, 5000], w_i uniform [0, 0.2], b uniform [2, 3] (seeded stream)."""
    from . import synth

    u = 0.5 * (synth.uniform(seed, 5 * count) + 1.0)
    N = np.exp(np.log(50.0) + u[0::5] * np.log(100.0))
    w = 0.2 * np.stack([u[1::5], u[2::5], u[3::5]], axis=1)
    b = 2.0 + u[4::5]
    return N, w, b


def solve_scaled(N, w, b, tol: float = 1e-9, check_every: int = 100, max_steps: int = 40000, verbose=True):
    """Ground states phi of the scaled problems (one batch class of 64^3 instances)."""
    k = len(N)
    n = [GRID] * k
    L = [2 * U_BOX] * k
    Nc = [4.0 * np.pi / 3.0] * k
    bt = capi.Batch(n, L, Nc)
    for i in range(k):
        R0, r0 = radius(N[i], b[i]), rho0(b[i])
        bt.upload(i, baseline(N[i], b[i]).astype(np.complex128))
        bt.set_lhy(i, R0 ** 2 * A_LHY * r0, R0 ** 2 * b[i] * r0 ** 1.5)
        bt.set_potential(i, potential_scaled(N[i], w[i], b[i]))
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    steps, en, ms = bt.ground_state(f, DT_FACTOR, tol, check_every, max_steps)
    phi = np.stack([bt.download(i).real.astype(np.float32) for i in range(k)])
    bt.close()
    if verbose:
        print(f"  solved {k} scaled ground states in {ms / 1e3:.1f} s, steps {int(np.abs(steps).min())}.."
              f"{int(np.abs(steps).max())}, converged {int((steps > 0).sum())}", flush=True)
    return phi, steps


def make_dataset(count: int, seed: int, chunk: int = 250):
    N, w, b = sample_params(count, seed)
    phis = []
    for s in range(0, count, chunk):
        phi, _ = solve_scaled(N[s:s + chunk], w[s:s + chunk], b[s:s + chunk])
        phis.append(phi)
    return N, w, b, np.concatenate(phis)


def train(N, w, b, phi, epochs: int = 12, batch: int = 4, lr: float = 1e-3, seed: int = 0, verbose=True):
    """Fits the network to the ground states phi (it learns phi - baseline); the
    features, baselines and targets of all samples stay in device memory."""
    k = len(N)
    P = GRID ** 3
    Xs = nn.DevBuf(k * P * nn.C0 * 2)
    bases = nn.DevBuf(k * P * 4)
    targets = nn.DevBuf(k * P * 4)
    L = nn._lib()
    import ctypes

    for i in range(k):
        X = features(N[i], w[i], b[i])
        h = nn.to_bf16(X)
        nn._ck(L.gp5_nn_upload(ctypes.c_void_p(Xs.ptr.value + i * h.nbytes), h.ctypes.data_as(ctypes.c_void_p),
                               h.nbytes))
        base = np.ascontiguousarray(X[..., 0])
        nn._ck(L.gp5_nn_upload(ctypes.c_void_p(bases.ptr.value + i * base.nbytes),
                               base.ctypes.data_as(ctypes.c_void_p), base.nbytes))
        t = np.ascontiguousarray(phi[i], dtype=np.float32)
        nn._ck(L.gp5_nn_upload(ctypes.c_void_p(targets.ptr.value + i * t.nbytes), t.ctypes.data_as(ctypes.c_void_p),
                               t.nbytes))
    net = nn.UNet(batch, GRID, seed=seed)
    rng = np.random.default_rng(seed)
    steps_total = epochs * (k // batch)
    hist = []
    t0 = time.time()
    it = 0
    for ep in range(epochs):
        order = rng.permutation(k)
        for j in range(k // batch):
            net.load_batch(Xs, bases, targets, order[j * batch:(j + 1) * batch])
            net.forward()
            net.backward(read_loss=False)  # losses summed on the device, read once per epoch
            net.step(0.5 * lr * (1.0 + np.cos(np.pi * it / steps_total)))  # cosine decay
            it += 1
        hist.append(net.take_loss_sum() / (k // batch))
        if verbose:
            print(f"  epoch {ep + 1}/{epochs}: mse {hist[-1]:.3e} ({time.time() - t0:.0f} s)", flush=True)
    return net.params(), hist


class Surrogate:
    """Inference with trained weights (float masters from train() or an .npz file)."""

    def __init__(self, params=None, path: str = DEFAULT_WEIGHTS, batch: int = 1):
        """batch > 1 also sets up a network of that batch size for predict_many /
        seed_many (one forward pass per `batch` droplets)."""
        if params is None:
            with np.load(path, allow_pickle=False) as z:
                params = {k: z[k] for k in z.files if "." in k}
        self.net = nn.UNet(1, GRID, params)
        self.batch_net = nn.UNet(int(batch), GRID, params) if batch > 1 else None

    @staticmethod
    def _set_features(net, N: float, w, b: float, index: int = 0):
        R0 = radius(N, b)
        a = [(wi / 0.2) ** 2 / 18.0 for wi in w]
        consts = [np.log(N / 50.0) / np.log(100.0), 2.0 * (b - 2.5), w[0] / 0.2, w[1] / 0.2, w[2] / 0.2]
        net.set_droplet_features(np.float32(R0 / SURFACE), np.float32(a), np.float32(consts), U_BOX, index)

    def predict(self, N: float, w, b: float, device_features: bool = True) -> np.ndarray:
        """phi on the scaled 64^3 grid.  The input channels are built on the device
        (gp5_nn_droplet_features, the float32 arithmetic of features()) unless
        device_features is false."""
        if device_features:
            self._set_features(self.net, N, w, b)
            self.net.forward()
            return self.net.output0()[0]
        self._X = features(N, w, b, out=getattr(self, "_X", None))
        X = self._X[None]
        self.net.set_batch(X, X[..., 0])
        self.net.forward()
        return self.net.output()[0]

    def predict_many(self, items):
        """predict() for a list of (N, w, b), `batch` droplets per forward pass when
        the Surrogate was made with batch > 1 (the kernels may then take another
        summation order than at batch 1: results agree to the bf16 level)."""
        if self.batch_net is None:
            return [self.predict(N, w, b) for N, w, b in items]
        out = []
        B = self.batch_net.B
        for c0 in range(0, len(items), B):
            chunk = items[c0:c0 + B]
            for j, (N, w, b) in enumerate(chunk):
                self._set_features(self.batch_net, N, w, b, j)
            self.batch_net.forward()  # samples past the chunk hold stale inputs; their outputs are unused
            y = self.batch_net.output0()
            out += [y[j] for j in range(len(chunk))]
        return out

    def seed_many(self, items):
        """seed() for a list of (N, w, b, n, L), the predictions batched (predict_many);
        the seeds come one at a time (a generator), so that only one is held."""
        phis = self.predict_many([(N, w, b) for N, w, b, _, _ in items])
        return (self._seed_from(phi, N, w, b, n, L) for phi, (N, w, b, n, L) in zip(phis, items))

    def seed(self, N: float, w, b: float, n: int, L: float) -> np.ndarray:
        """psi on the cubic grid n^3 over [-L/2, L/2]^3 (complex128, column-major).

        The prediction is projected onto the symmetry of the problem before it is
        used: imaginary time removes a shape error at the rate of the softest mode
        that the error contains, and the soft surface deformations of a droplet
        (quadrupole first) are absent from a symmetric seed.  Untrapped droplets
        (w = 0) are spherical: the seed is the shell average of the prediction as a
        radial profile.  Otherwise the prediction is averaged over the axis
        reflections and over the permutations of axes with equal trap frequencies.
        The profile is clipped at zero and tapered to zero between |u| = 2 and 2.6,
        since stray amplitude far from the droplet is a dilute gas that imaginary
        time removes only slowly."""
        return self._seed_from(self.predict(N, w, b), N, w, b, n, L)

    def _seed_from(self, phi: np.ndarray, N: float, w, b: float, n: int, L: float) -> np.ndarray:
        """The seed of seed() from a prediction phi."""
        phi = phi.astype(np.float64)
        R0 = radius(N, b)
        x = -0.5 * L + (np.arange(n) + 0.5) * (L / n)
        if max(abs(float(v)) for v in w) == 0.0:
            ru, prof = radial_profile(phi)
            fp = np.maximum(prof, 0.0) * _taper(ru)
            if n % 2 == 0:
                # the cell centres are mirror-symmetric about 0: one octant, then flips
                xa = x[n // 2:]
                rx = np.sqrt(xa[:, None, None] ** 2 + xa[None, :, None] ** 2 + xa[None, None, :] ** 2) / R0
                out = np.interp(rx, ru, fp, right=0.0)
                for ax in range(3):
                    out = np.concatenate([np.flip(out, axis=ax), out], axis=ax)
            else:
                rx = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2) / R0
                out = np.interp(rx, ru, fp, right=0.0)
            # the field is radial, so its transpose (Fortran-ordered, no copy) holds the
            # same values up to the summation order of x^2 + y^2 + z^2; the solver takes
            # column-major fields, so the upload then needs no transposing copy
            return (np.sqrt(rho0(b)) * out).astype(np.complex128).T
        phi = symmetrize(phi, w)
        _, r = _grids(GRID)
        return sqrt_rho_interp(np.maximum(phi, 0.0) * _taper(r), N, b, n, L)


def _taper(r):
    return np.where(r <= 2.0, 1.0, np.where(r >= 2.6, 0.0, np.cos(0.5 * np.pi * (r - 2.0) / 0.6) ** 2))


_SHELLS = {}


def _shells(n: int, dr: float):
    """Shell index of every cell of the scaled n^3 grid, the cells per shell, the
    occupied shells and their mean radii (they depend on the grid only)."""
    key = (n, dr)
    if key not in _SHELLS:
        _, r = _grids(n)
        idx = np.floor(r / dr).astype(int).ravel()
        cnt = np.bincount(idx)
        rs = np.bincount(idx, weights=r.ravel())
        keep = cnt > 0
        _SHELLS[key] = (idx, cnt, keep, rs[keep] / cnt[keep])
    return _SHELLS[key]


def radial_profile(phi: np.ndarray, dr: float = None):
    """Shell averages of a field on the scaled grid: (shell radii, mean values); the
    first shell is the centre."""
    h = 2 * U_BOX / phi.shape[0]
    dr = dr or 0.5 * h
    idx, cnt, keep, rmean = _shells(phi.shape[0], dr)
    sums = np.bincount(idx, weights=phi.ravel(), minlength=cnt.size)
    return rmean, sums[keep] / cnt[keep]


def symmetrize(phi: np.ndarray, w) -> np.ndarray:
    """Average over the axis reflections and the permutations of axes with equal w."""
    out = np.zeros_like(phi)
    k = 0
    for fx in (False, True):
        for fy in (False, True):
            for fz in (False, True):
                a = phi[::-1] if fx else phi
                a = a[:, ::-1] if fy else a
                a = a[:, :, ::-1] if fz else a
                out += a
                k += 1
    out /= k
    perms = [p for p in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
             if all(abs(float(w[p[i]]) - float(w[i])) < 1e-12 for i in range(3))]
    return sum(np.transpose(out, p) for p in perms) / len(perms)


def sqrt_rho_interp(phi: np.ndarray, N: float, b: float, n: int, L: float) -> np.ndarray:
    """sqrt(rho0) phi(x / R0) at the cell centres of the instance grid (trilinear, zero outside)."""
    R0 = radius(N, b)
    x = -0.5 * L + (np.arange(n) + 0.5) * (L / n)
    # fractional index of u = x / R0 on the scaled grid
    g = (x / R0 + U_BOX) / (2 * U_BOX / GRID) - 0.5
    i0 = np.floor(g).astype(int)
    t = g - i0
    P = np.pad(phi, 1)  # zero outside the scaled box
    i0 = np.clip(i0 + 1, 0, GRID)
    i1 = np.clip(i0 + 1, 0, GRID + 1)
    out = np.zeros((n, n, n))
    for az, iz, tz in ((0, i0, 1 - t), (1, i1, t)):
        for ay, iy, ty in ((0, i0, 1 - t), (1, i1, t)):
            for ax, ix, tx in ((0, i0, 1 - t), (1, i1, t)):
                out += P[np.ix_(ix, iy, iz)] * (tx[:, None, None] * ty[None, :, None] * tz[None, None, :])
    inside = np.abs(x / R0) <= U_BOX
    out *= inside[:, None, None] & inside[None, :, None] & inside[None, None, :]
    return (np.sqrt(rho0(b)) * out).astype(np.complex128)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--count", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)
    t0 = time.time()
    print(f"generating {args.count} ground states", flush=True)
    N, w, b, phi = make_dataset(args.count, args.seed)
    print(f"dataset ready ({time.time() - t0:.0f} s); training", flush=True)
    params, hist = train(N, w, b, phi, epochs=args.epochs)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.savez_compressed(args.out, loss=np.array(hist), **params)
    print(f"saved {args.out} ({time.time() - t0:.0f} s)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
