"""ctypes binding of the C ABI in include/gp5.h (libgp5.so).

This is the thinnest host layer: it exposes the C entry points one to one,
with numpy arrays in the reference layout (complex128, shape (Nx,Ny,Nz,nC),
Fortran order; src/traits.h:12-20) on the host side.  The reference-shaped
API (geometry / field / LHY / stepper, the gpCpp module) sits on top.

Loading fails loudly when libgp5.so has not been built: there is no CPU
fallback for any operation.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GP5_LIB_PATH selects another build of the same ABI (A/B timing of build options)
LIB_PATH = os.environ.get("GP5_LIB_PATH") or os.path.join(_HERE, "libgp5.so")

GP5_OK = 0
GP5_ERR_ARG = 1
GP5_ERR_HIP = 2
GP5_ERR_UNSUPPORTED = 3
GP5_ERR_ALLOC = 4
GP5_ERR_NONFINITE = 5

GP5_FORWARD, GP5_BACKWARD = 0, 1
GP5_FUNC_GP, GP5_FUNC_LHY, GP5_FUNC_LHY_UNLOCKED = 0, 1, 2
GP5_STEP_EULERO, GP5_STEP_RK4 = 0, 1

# every symbol include/gp5.h declares
EXPORTS = [
    "gp5_last_error", "gp5_version", "gp5_ctx_create", "gp5_ctx_destroy", "gp5_ctx_sync",
    "gp5_ctx_points", "gp5_ctx_fft_plans", "gp5_ctx_set_fft_plans", "gp5_ctx_fft_plan_name", "gp5_buf_alloc", "gp5_buf_free", "gp5_buf_size", "gp5_buf_upload",
    "gp5_buf_download", "gp5_buf_copy", "gp5_buf_zero", "gp5_fft", "gp5_laplacian",
    "gp5_functional_apply", "gp5_norm", "gp5_normalize", "gp5_max_density", "gp5_energy",
    "gp5_phase_constraint", "gp5_axpy", "gp5_rk4_combine", "gp5_advance", "gp5_advance_noalloc", "gp5_advance_timed", "gp5_profile_advance",
    "gp5_kernel_kind_name", "gp5_resample", "gp5_decompose", "gp5_h5_write", "gp5_h5_read", "gp5_h5_shape",
    "gp5_batch_create", "gp5_batch_destroy", "gp5_batch_classes", "gp5_batch_class_of",
    "gp5_batch_upload", "gp5_batch_download", "gp5_batch_set_lhy", "gp5_batch_set_potential",
    "gp5_batch_advance", "gp5_batch_tune", "gp5_batch_ground_state", "gp5_batch_resample",
]
GP5_NKINDS = 8


class GP5Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class FunctionalDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("has_couplings", ctypes.c_int),
        ("couplings", ctypes.c_double * 4), ("masses", ctypes.c_double * 2),
        ("lhy_a", ctypes.c_double), ("lhy_b", ctypes.c_double),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("eta", ctypes.c_double),
        ("potential", ctypes.c_void_p),
    ]


class StepperDesc(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("tau_re", ctypes.c_double), ("tau_im", ctypes.c_double),
        ("renormalize", ctypes.c_int), ("normalizations", ctypes.c_double * 2),
        ("phase", ctypes.c_void_p),
    ]


_lib = None


def load() -> ctypes.CDLL:
    """Load libgp5.so; raises if it is missing (build it with python -m gp5_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m gp5_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    P, I, D, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_size_t
    PP = ctypes.POINTER(ctypes.c_void_p)
    DP = ctypes.POINTER(ctypes.c_double)
    sig = {
        "gp5_last_error": ([], ctypes.c_char_p),
        "gp5_version": ([], I),
        "gp5_ctx_create": ([PP, ctypes.POINTER(ctypes.c_int), DP, DP, I], I),
        "gp5_ctx_destroy": ([P], I),
        "gp5_ctx_sync": ([P], I),
        "gp5_ctx_points": ([P], S),
        "gp5_ctx_fft_plans": ([P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], I),
        "gp5_ctx_set_fft_plans": ([P, ctypes.POINTER(ctypes.c_int)], I),
        "gp5_ctx_fft_plan_name": ([P, I], ctypes.c_char_p),
        "gp5_buf_alloc": ([P, S, PP], I),
        "gp5_buf_free": ([P], I),
        "gp5_buf_size": ([P], S),
        "gp5_buf_upload": ([P, P, S], I),
        "gp5_buf_download": ([P, P, S], I),
        "gp5_buf_copy": ([P, P], I),
        "gp5_buf_zero": ([P], I),
        "gp5_fft": ([P, P, P, I], I),
        "gp5_laplacian": ([P, P, P], I),
        "gp5_functional_apply": ([P, ctypes.POINTER(FunctionalDesc), P, P, D], I),
        "gp5_norm": ([P, P, I, DP], I),
        "gp5_normalize": ([P, P, I, D], I),
        "gp5_max_density": ([P, P, I, DP], I),
        "gp5_energy": ([P, ctypes.POINTER(FunctionalDesc), P, DP], I),
        "gp5_phase_constraint": ([P, P, P], I),
        "gp5_axpy": ([P, P, P, D, D, P], I),
        "gp5_rk4_combine": ([P, P, P, D, D, P, P, P, P], I),
        "gp5_advance": ([P, ctypes.POINTER(FunctionalDesc), ctypes.POINTER(StepperDesc), P, I, P, P], I),
        "gp5_advance_noalloc": ([P, ctypes.POINTER(FunctionalDesc), ctypes.POINTER(StepperDesc), P, I], I),
        "gp5_advance_timed": ([P, ctypes.POINTER(FunctionalDesc), ctypes.POINTER(StepperDesc), P, I, DP], I),
        "gp5_profile_advance": ([P, ctypes.POINTER(FunctionalDesc), ctypes.POINTER(StepperDesc), P, I, DP,
                                 ctypes.POINTER(ctypes.c_int)], I),
        "gp5_kernel_kind_name": ([I], ctypes.c_char_p),
        "gp5_resample": ([P, P, P, P], I),
        "gp5_decompose": ([P, P, I, D, P, ctypes.POINTER(ctypes.c_int)], I),
        "gp5_h5_write": ([ctypes.c_char_p, P, ctypes.POINTER(ctypes.c_int), I], I),
        "gp5_h5_read": ([ctypes.c_char_p, P, ctypes.POINTER(ctypes.c_int), I], I),
        "gp5_h5_shape": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)], I),
        "gp5_batch_create": ([PP, I, ctypes.POINTER(ctypes.c_int), DP, DP], I),
        "gp5_batch_destroy": ([P], I),
        "gp5_batch_classes": ([P], I),
        "gp5_batch_class_of": ([P, I], I),
        "gp5_batch_upload": ([P, I, P], I),
        "gp5_batch_download": ([P, I, P], I),
        "gp5_batch_set_lhy": ([P, I, D, D], I),
        "gp5_batch_set_potential": ([P, I, P], I),
        "gp5_batch_advance": ([P, ctypes.POINTER(FunctionalDesc), D, I], I),
        "gp5_batch_tune": ([P, ctypes.POINTER(FunctionalDesc), D, I, ctypes.c_char_p, S], I),
        "gp5_batch_ground_state": ([P, ctypes.POINTER(FunctionalDesc), D, D, I, ctypes.c_longlong,
                                    ctypes.POINTER(ctypes.c_longlong), DP, DP], I),
        "gp5_batch_resample": ([P, I, P, I], I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != GP5_OK:
        msg = load().gp5_last_error().decode(errors="replace")
        raise GP5Error(rc, msg)


class Context:
    """gp5_ctx: grid, FFT plans, stream and scratch of one field shape."""

    def __init__(self, shape: Sequence[int], left: Sequence[float], right: Sequence[float], nc: int = 1):
        L = load()
        self.shape = tuple(int(s) for s in shape)
        self.left = tuple(float(x) for x in left)
        self.right = tuple(float(x) for x in right)
        self.nc = int(nc)
        h = ctypes.c_void_p()
        sh = (ctypes.c_int * 3)(*self.shape)
        lf = (ctypes.c_double * 3)(*self.left)
        rt = (ctypes.c_double * 3)(*self.right)
        check(L.gp5_ctx_create(ctypes.byref(h), sh, lf, rt, self.nc))
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def M(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    def close(self) -> None:
        if getattr(self, "_h", None):
            check(load().gp5_ctx_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                load().gp5_ctx_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def field(self, data: Optional[np.ndarray] = None) -> "Buffer":
        b = Buffer(self, 2 * self.M * self.nc)
        if data is not None:
            b.upload_complex(data)
        return b

    def real(self, data: Optional[np.ndarray] = None) -> "Buffer":
        b = Buffer(self, self.M * self.nc)
        if data is not None:
            b.upload_real(data)
        return b

    def fft_plans(self):
        """(variants, counts) of the register plans per role (x lines, y columns, z pass)."""
        v, c = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        check(load().gp5_ctx_fft_plans(self.handle, v, c))
        return list(v), list(c)

    def set_fft_plans(self, variants) -> None:
        check(load().gp5_ctx_set_fft_plans(self.handle, (ctypes.c_int * 3)(*[int(x) for x in variants])))

    def fft_plan_names(self):
        return [load().gp5_ctx_fft_plan_name(self.handle, r).decode() for r in range(3)]

    def sync(self) -> None:
        check(load().gp5_ctx_sync(self._h))


class Buffer:
    """gp5_buf: device array of doubles owned by a context."""

    def __init__(self, ctx: Context, n_doubles: int):
        self.ctx = ctx
        self.n = int(n_doubles)
        h = ctypes.c_void_p()
        check(load().gp5_buf_alloc(ctx.handle, self.n, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def free(self) -> None:
        if getattr(self, "_h", None):
            check(load().gp5_buf_free(self._h))
            self._h = None

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                load().gp5_buf_free(self._h)
                self._h = None
        except Exception:
            pass

    def _shape(self):
        s = self.ctx.shape
        return (s[0], s[1], s[2], self.ctx.nc)

    def upload_complex(self, a: np.ndarray) -> None:
        a = np.asarray(a)
        if a.ndim == 3 and self.ctx.nc == 1:
            a = a[..., None]
        if a.shape != self._shape():
            raise GP5Error(GP5_ERR_ARG, f"array of shape {a.shape} for a field of shape {self._shape()}")
        flat = np.ascontiguousarray(np.asarray(a, dtype=np.complex128).ravel(order="F"))
        check(load().gp5_buf_upload(self._h, flat.ctypes.data_as(ctypes.c_void_p), self.n))

    def upload_real(self, a: np.ndarray) -> None:
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 3 and self.ctx.nc == 1:
            a = a[..., None]
        if a.shape != self._shape():
            raise GP5Error(GP5_ERR_ARG, f"array of shape {a.shape} for a real field of shape {self._shape()}")
        flat = np.ascontiguousarray(a.ravel(order="F"))
        check(load().gp5_buf_upload(self._h, flat.ctypes.data_as(ctypes.c_void_p), self.n))

    def download_complex(self) -> np.ndarray:
        flat = np.empty(self.n // 2, dtype=np.complex128)
        check(load().gp5_buf_download(self._h, flat.ctypes.data_as(ctypes.c_void_p), self.n))
        return flat.reshape(self._shape(), order="F")


def functional_desc(kind: int = GP5_FUNC_LHY, *, couplings=None, masses=(1.0, 1.0), lhy_a: float = -3.0,
                    lhy_b: float = 2.5, alpha: float = 0.0, beta: float = 0.0, eta: float = 0.0,
                    potential: Optional[Buffer] = None) -> FunctionalDesc:
    f = FunctionalDesc()
    f.kind = kind
    f.has_couplings = 0 if couplings is None else 1
    if couplings is not None:
        g = np.asarray(couplings, dtype=float)
        for i in range(g.shape[0]):
            for j in range(g.shape[1]):
                f.couplings[i + 2 * j] = g[i, j]
    f.masses[0] = masses[0]
    f.masses[1] = masses[1] if len(masses) > 1 else 1.0
    f.lhy_a, f.lhy_b = lhy_a, lhy_b
    f.alpha, f.beta, f.eta = alpha, beta, eta
    f.potential = potential.handle.value if potential is not None else None
    return f


def stepper_desc(kind: int = GP5_STEP_RK4, *, dt: float, imaginary: bool, renormalize=None,
                 phase: Optional[Buffer] = None) -> StepperDesc:
    s = StepperDesc()
    s.kind = kind
    s.tau_re, s.tau_im = (dt, 0.0) if imaginary else (0.0, dt)
    s.renormalize = 0 if renormalize is None else 1
    if renormalize is not None:
        for c, N in enumerate(renormalize):
            s.normalizations[c] = float(N)
    s.phase = phase.handle.value if phase is not None else None
    return s


def fft(ctx: Context, src: Buffer, dst: Buffer, direction: int) -> None:
    check(load().gp5_fft(ctx.handle, src.handle, dst.handle, int(direction)))


def laplacian(ctx: Context, src: Buffer, dst: Buffer) -> None:
    check(load().gp5_laplacian(ctx.handle, src.handle, dst.handle))


def functional_apply(ctx: Context, f: FunctionalDesc, src: Buffer, dst: Buffer, time: float = 0.0) -> None:
    check(load().gp5_functional_apply(ctx.handle, ctypes.byref(f), src.handle, dst.handle, float(time)))


def norm(ctx: Context, f: Buffer, c: int = 0) -> float:
    out = ctypes.c_double()
    check(load().gp5_norm(ctx.handle, f.handle, int(c), ctypes.byref(out)))
    return out.value


def normalize(ctx: Context, f: Buffer, c: int, N: float) -> None:
    check(load().gp5_normalize(ctx.handle, f.handle, int(c), float(N)))


def max_density(ctx: Context, f: Buffer, c: int = 0) -> float:
    out = ctypes.c_double()
    check(load().gp5_max_density(ctx.handle, f.handle, int(c), ctypes.byref(out)))
    return out.value


def energy(ctx: Context, f: FunctionalDesc, psi: Buffer) -> float:
    out = ctypes.c_double()
    check(load().gp5_energy(ctx.handle, ctypes.byref(f), psi.handle, ctypes.byref(out)))
    return out.value


def phase_constraint(ctx: Context, psi: Buffer, phase: Buffer) -> None:
    check(load().gp5_phase_constraint(ctx.handle, psi.handle, phase.handle))


def axpy(ctx: Context, out: Buffer, a: Buffer, c: complex, k: Buffer) -> None:
    """out = a - c k (one stage update of src/stepper.cpp:30-70 as its own pass)."""
    check(load().gp5_axpy(ctx.handle, out.handle, a.handle, float(np.real(c)), float(np.imag(c)), k.handle))


def rk4_combine(ctx: Context, out: Buffer, a: Buffer, c: complex, k1: Buffer, k2: Buffer, k3: Buffer,
                k4: Buffer) -> None:
    """out = a - c (((k1 + 2 k2) + 2 k3) + k4) (src/stepper.cpp:64)."""
    check(load().gp5_rk4_combine(ctx.handle, out.handle, a.handle, float(np.real(c)), float(np.imag(c)), k1.handle,
                                 k2.handle, k3.handle, k4.handle))


def advance(ctx: Context, f: FunctionalDesc, s: StepperDesc, psi: Buffer, nsteps: int, *,
            log: bool = False):
    """Advance psi in place; with log=True returns (norms[nsteps, nC], energies[nsteps])."""
    L = load()
    if log:
        norms = np.zeros((int(nsteps), ctx.nc), dtype=np.float64)
        ens = np.zeros(int(nsteps), dtype=np.float64)
        check(L.gp5_advance(ctx.handle, ctypes.byref(f), ctypes.byref(s), psi.handle, int(nsteps),
                            norms.ctypes.data_as(ctypes.c_void_p), ens.ctypes.data_as(ctypes.c_void_p)))
        return norms, ens
    check(L.gp5_advance(ctx.handle, ctypes.byref(f), ctypes.byref(s), psi.handle, int(nsteps), None, None))
    return None


def advance_noalloc(ctx: Context, f: FunctionalDesc, s: StepperDesc, psi: Buffer, nsteps: int) -> None:
    """gp5_advance_noalloc: advance on the library's static device scratch (no allocation)."""
    check(load().gp5_advance_noalloc(ctx.handle, ctypes.byref(f), ctypes.byref(s), psi.handle, int(nsteps)))


def advance_timed(ctx: Context, f: FunctionalDesc, s: StepperDesc, psi: Buffer, nsteps: int) -> float:
    """Run nsteps and return the device time between HIP events on the context stream (ms)."""
    ms = ctypes.c_double()
    check(load().gp5_advance_timed(ctx.handle, ctypes.byref(f), ctypes.byref(s), psi.handle, int(nsteps),
                                   ctypes.byref(ms)))
    return ms.value


def profile_advance(ctx: Context, f: FunctionalDesc, s: StepperDesc, psi: Buffer, nsteps: int) -> dict:
    """Per-kernel-kind device time of nsteps (HIP events around each launch on the context stream).
    Returns {kind_name: (total_ms, launches)} for kinds that ran."""
    L = load()
    ms = (ctypes.c_double * GP5_NKINDS)()
    cnt = (ctypes.c_int * GP5_NKINDS)()
    check(L.gp5_profile_advance(ctx.handle, ctypes.byref(f), ctypes.byref(s), psi.handle, int(nsteps), ms, cnt))
    out = {}
    for k in range(GP5_NKINDS):
        if cnt[k]:
            out[L.gp5_kernel_kind_name(k).decode()] = (ms[k], cnt[k])
    return out


def resample(src_ctx: Context, src: Buffer, dst_ctx: Context, dst: Buffer) -> None:
    """dst = src carried to dst's grid by trigonometric interpolation (same box)."""
    check(load().gp5_resample(src_ctx.handle, src.handle, dst_ctx.handle, dst.handle))


def decompose(ctx: Context, psi: Buffer, cutoff: float, comp: int = 0):
    """Droplet labels (src/cluster.cpp:17-82) on the GPU: an int32 (Nx, Ny, Nz, nC)
    Fortran-ordered array (-1 below the cut-off, droplets 1..n, other
    components 0) and n."""
    labels = np.zeros(ctx.shape + (ctx.nc,), dtype=np.int32, order="F")
    n = ctypes.c_int()
    check(load().gp5_decompose(ctx.handle, psi.handle, int(comp), float(cutoff), labels.ctypes.data,
                               ctypes.byref(n)))
    return labels, n.value


def h5_shape(filename: str):
    """((Nx, Ny, Nz), nC) of the "field" dataset of an HDF5 field file."""
    shape, nc = (ctypes.c_int * 3)(), ctypes.c_int()
    check(load().gp5_h5_shape(os.fsencode(filename), shape, ctypes.byref(nc)))
    return tuple(shape), nc.value


def h5_write(filename: str, array) -> None:
    """Write a (Nx, Ny, Nz, nC) complex array as the reference's HDF5 "field"
    dataset (src/io.cpp:56-152); a 3-D array is one component."""
    a = np.asarray(array)
    if a.ndim == 3:
        a = a[..., None]
    if a.ndim != 4:
        raise ValueError("expected a (Nx, Ny, Nz, nC) array")
    a = np.asfortranarray(a, dtype=np.complex128)
    shape = (ctypes.c_int * 3)(*a.shape[:3])
    check(load().gp5_h5_write(os.fsencode(filename), a.ctypes.data, shape, a.shape[3]))


def h5_read(filename: str, shape: Optional[Sequence[int]] = None, nc: Optional[int] = None) -> np.ndarray:
    """Read an HDF5 field file (src/io.cpp:154-254) as a (Nx, Ny, Nz, nC)
    complex128 Fortran-ordered array; with shape/nc given, a file of other
    dims raises."""
    fshape, fnc = h5_shape(filename)
    shape = tuple(fshape if shape is None else shape)
    nc = fnc if nc is None else nc
    out = np.zeros(tuple(shape) + (nc,), dtype=np.complex128, order="F")
    check(load().gp5_h5_read(os.fsencode(filename), out.ctypes.data, (ctypes.c_int * 3)(*shape), nc))
    return out


class Batch:
    """gp5_batch: many single-component instances (cube grid n_i on [-L_i/2, L_i/2]^3,
    normalisation N_i) stepped together, grouped by (n, L) into one context each."""

    def __init__(self, n, L, N):
        L_ = load()
        self.n = [int(x) for x in n]
        self.L = [float(x) for x in L]
        self.N = [float(x) for x in N]
        k = len(self.n)
        if not (len(self.L) == k and len(self.N) == k):
            raise GP5Error(GP5_ERR_ARG, "n, L and N need one entry per instance")
        h = ctypes.c_void_p()
        check(L_.gp5_batch_create(ctypes.byref(h), k, (ctypes.c_int * k)(*self.n), (ctypes.c_double * k)(*self.L),
                                  (ctypes.c_double * k)(*self.N)))
        self.handle = h
        self.count = k

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(load().gp5_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def classes(self) -> int:
        return load().gp5_batch_classes(self.handle)

    def class_of(self, i: int) -> int:
        return load().gp5_batch_class_of(self.handle, i)

    def upload(self, i: int, a: np.ndarray) -> None:
        a = np.asarray(a, dtype=np.complex128)
        n = self.n[i]
        if a.ndim == 4:
            a = a[..., 0]
        if a.shape != (n, n, n):
            raise GP5Error(GP5_ERR_ARG, f"instance {i} holds a {n}^3 field, got {a.shape}")
        flat = np.ascontiguousarray(a.ravel(order="F"))
        check(load().gp5_batch_upload(self.handle, i, flat.ctypes.data_as(ctypes.c_void_p)))

    def download(self, i: int) -> np.ndarray:
        n = self.n[i]
        flat = np.empty(n ** 3, dtype=np.complex128)
        check(load().gp5_batch_download(self.handle, i, flat.ctypes.data_as(ctypes.c_void_p)))
        return flat.reshape((n, n, n), order="F")

    def set_lhy(self, i: int, a: float, b: float) -> None:
        check(load().gp5_batch_set_lhy(self.handle, i, float(a), float(b)))

    def set_potential(self, i: int, V: np.ndarray) -> None:
        V = np.asarray(V, dtype=np.float64)
        n = self.n[i]
        if V.ndim == 4:
            V = V[..., 0]
        if V.shape != (n, n, n):
            raise GP5Error(GP5_ERR_ARG, f"instance {i} holds a {n}^3 potential, got {V.shape}")
        flat = np.ascontiguousarray(V.ravel(order="F"))
        check(load().gp5_batch_set_potential(self.handle, i, flat.ctypes.data_as(ctypes.c_void_p)))

    def advance(self, f: FunctionalDesc, dt_factor: float, nsteps: int) -> None:
        check(load().gp5_batch_advance(self.handle, ctypes.byref(f), float(dt_factor), int(nsteps)))

    def tune(self, f: FunctionalDesc, dt_factor: float, steps: int = 2) -> str:
        buf = ctypes.create_string_buffer(8192)
        check(load().gp5_batch_tune(self.handle, ctypes.byref(f), float(dt_factor), int(steps), buf, len(buf)))
        return buf.value.decode()

    def resample_from(self, i: int, src: "Batch", j: int) -> None:
        """Instance i = instance j of src resampled to this instance's grid (same box)."""
        check(load().gp5_batch_resample(self.handle, int(i), src.handle, int(j)))

    def ground_state(self, f: FunctionalDesc, dt_factor: float, tol: float, check_every: int = 100,
                     max_steps: int = 0):
        """Returns (steps[count], energy[count], elapsed_ms); steps < 0: not converged."""
        k = self.count
        steps = (ctypes.c_longlong * k)()
        en = (ctypes.c_double * k)()
        ms = ctypes.c_double()
        check(load().gp5_batch_ground_state(self.handle, ctypes.byref(f), float(dt_factor), float(tol),
                                            int(check_every), int(max_steps), steps, en, ctypes.byref(ms)))
        return np.array(steps[:], dtype=np.int64), np.array(en[:]), ms.value
