"""ctypes front end of the CPU oracle (gp5_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (gp5_amd) never
imports it.

Arrays follow the reference layout (src/traits.h:12-20): complex128, shape
(Nx, Ny, Nz, nC), Fortran (column-major) order, so ``arr.ravel(order="F")``
is the interleaved buffer the C code reads.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

OR_GP, OR_LHY, OR_UNLOCKED = 0, 1, 2


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


class _Geom(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("left", ctypes.c_double * 3),
                ("right", ctypes.c_double * 3), ("nc", ctypes.c_int)]


class _Func(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("has_couplings", ctypes.c_int),
                ("g", ctypes.c_double * 4), ("inv_mass", ctypes.c_double * 2),
                ("lhy_a", ctypes.c_double), ("lhy_b", ctypes.c_double),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("eta", ctypes.c_double), ("V", ctypes.c_void_p)]


class _Stepper(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("tau_re", ctypes.c_double),
                ("tau_im", ctypes.c_double), ("renorm", ctypes.c_int),
                ("N", ctypes.c_double * 2), ("phase", ctypes.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        D = ctypes.c_double
        L.or_fft3d.argtypes = [P, P, I, I, I, I, I]
        L.or_fft3d.restype = I
        L.or_laplacian.argtypes = [P, P, ctypes.POINTER(_Geom), P]
        L.or_laplacian.restype = I
        L.or_apply.argtypes = [ctypes.POINTER(_Func), P, P, ctypes.POINTER(_Geom), P]
        L.or_apply.restype = I
        L.or_energy.argtypes = [ctypes.POINTER(_Func), P, ctypes.POINTER(_Geom), P, P]
        L.or_energy.restype = D
        L.or_norm.argtypes = [P, ctypes.POINTER(_Geom), I]
        L.or_norm.restype = D
        L.or_normalize.argtypes = [P, ctypes.POINTER(_Geom), I, D]
        L.or_normalize.restype = None
        L.or_advance.argtypes = [ctypes.POINTER(_Func), ctypes.POINTER(_Stepper), P,
                                 ctypes.POINTER(_Geom), I]
        L.or_advance.restype = I
        L.or_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, P]
        L.or_uniform.restype = None
        L.or_momentum.argtypes = [I, I, D]
        L.or_momentum.restype = D
        L.or_num_threads.argtypes = []
        L.or_num_threads.restype = I
        L.or_set_threads.argtypes = [I]
        L.or_set_threads.restype = None
        L.or_decompose.argtypes = [P, I, I, I, I, I, D, P]
        L.or_decompose.restype = I
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().or_set_threads(int(n))


def num_threads() -> int:
    return lib().or_num_threads()


def _f(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.complex128 or not a.flags.f_contiguous:
        a = np.asfortranarray(a, dtype=np.complex128)
    return a


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Geometry:
    shape: Sequence[int]
    left: Sequence[float]
    right: Sequence[float]
    nc: int = 1

    def c(self) -> _Geom:
        g = _Geom()
        for d in range(3):
            g.n[d] = int(self.shape[d])
            g.left[d] = float(self.left[d])
            g.right[d] = float(self.right[d])
        g.nc = int(self.nc)
        return g

    @property
    def dx(self):
        return [(self.right[d] - self.left[d]) / self.shape[d] for d in range(3)]

    @property
    def dV(self) -> float:
        dx = self.dx
        return dx[0] * dx[1] * dx[2]

    def axes(self):
        """Cell-centred coordinates (src/geometry.cpp:165-188)."""
        return [self.left[d] + (np.arange(self.shape[d]) + 0.5) * self.dx[d] for d in range(3)]

    def grid(self):
        return np.meshgrid(*self.axes(), indexing="ij")


@dataclass
class Functional:
    kind: int = OR_LHY
    couplings: Optional[np.ndarray] = None
    masses: Sequence[float] = (1.0, 1.0)
    lhy_a: float = -3.0
    lhy_b: float = 2.5
    alpha: float = 0.0
    beta: float = 0.0
    eta: float = 0.0
    V: Optional[np.ndarray] = None  # real, shape (Nx,Ny,Nz,nC)
    _keep: list = field(default_factory=list)

    def c(self) -> _Func:
        f = _Func()
        f.kind = self.kind
        f.has_couplings = 0 if self.couplings is None else 1
        if self.couplings is not None:
            g = np.asarray(self.couplings, dtype=float)
            nc = g.shape[0]
            for i in range(nc):
                for j in range(nc):
                    f.g[i + 2 * j] = g[i, j]
        f.inv_mass[0] = 1.0 / self.masses[0]
        f.inv_mass[1] = 1.0 / (self.masses[1] if len(self.masses) > 1 else 1.0)
        f.lhy_a, f.lhy_b = self.lhy_a, self.lhy_b
        f.alpha, f.beta, f.eta = self.alpha, self.beta, self.eta
        if self.V is not None:
            V = np.asfortranarray(self.V, dtype=np.float64)
            self._keep = [V]
            f.V = V.ctypes.data
        else:
            f.V = None
        return f


def fft3d(a: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalised 3-D DFT of each component (src/fourierTransform.cpp:184-214)."""
    a = _f(a)
    if a.ndim == 3:
        a = a[..., None]
    out = np.empty_like(a, order="F")
    nx, ny, nz, nc = a.shape
    rc = lib().or_fft3d(_ptr(a), _ptr(out), nx, ny, nz, nc, int(sign))
    assert rc == 0
    return out


def laplacian(a: np.ndarray, geo: Geometry) -> np.ndarray:
    a = _f(a)
    out = np.empty_like(a, order="F")
    work = np.empty_like(a, order="F")
    g = geo.c()
    assert lib().or_laplacian(_ptr(a), _ptr(out), ctypes.byref(g), _ptr(work)) == 0
    return out


def apply(func: Functional, a: np.ndarray, geo: Geometry) -> np.ndarray:
    a = _f(a)
    out = np.empty_like(a, order="F")
    work = np.empty_like(a, order="F")
    g = geo.c()
    f = func.c()
    rc = lib().or_apply(ctypes.byref(f), _ptr(a), _ptr(out), ctypes.byref(g), _ptr(work))
    if rc != 0:
        raise RuntimeError(f"or_apply failed ({rc})")
    return out


def energy(func: Functional, a: np.ndarray, geo: Geometry) -> float:
    a = _f(a)
    work = np.empty_like(a, order="F")
    lap = np.empty_like(a, order="F")
    g = geo.c()
    f = func.c()
    return lib().or_energy(ctypes.byref(f), _ptr(a), ctypes.byref(g), _ptr(work), _ptr(lap))


def norm(a: np.ndarray, geo: Geometry, c: int = 0) -> float:
    a = _f(a)
    g = geo.c()
    return lib().or_norm(_ptr(a), ctypes.byref(g), int(c))


def normalize(a: np.ndarray, geo: Geometry, c: int, N: float) -> np.ndarray:
    a = _f(a).copy(order="F")
    g = geo.c()
    lib().or_normalize(_ptr(a), ctypes.byref(g), int(c), float(N))
    return a


def advance(func: Functional, a: np.ndarray, geo: Geometry, *, stepper: str = "RK4",
            dt: float = 1e-3, imaginary: bool = True, nsteps: int = 1,
            renormalize: Optional[Sequence[float]] = None,
            phase: Optional[np.ndarray] = None) -> np.ndarray:
    """timeStepper(dt, name, isImaginary).advance(field, nsteps) (src/gpCpp.cpp:202-269)."""
    a = _f(a).copy(order="F")
    s = _Stepper()
    if stepper == "eulero":
        s.kind = 0
    elif stepper == "RK4":
        s.kind = 1
    else:
        raise RuntimeError(f"Unknown stepper {stepper}")
    s.tau_re, s.tau_im = (dt, 0.0) if imaginary else (0.0, dt)
    s.renorm = 0 if renormalize is None else 1
    if renormalize is not None:
        for c, N in enumerate(renormalize):
            s.N[c] = float(N)
    keep = None
    if phase is not None:
        keep = np.asfortranarray(phase, dtype=np.float64)
        s.phase = keep.ctypes.data
    g = geo.c()
    f = func.c()
    rc = lib().or_advance(ctypes.byref(f), ctypes.byref(s), _ptr(a), ctypes.byref(g), int(nsteps))
    if rc != 0:
        raise RuntimeError(f"or_advance failed ({rc})")
    return a


def decompose(a: np.ndarray, cutoff: float, component: int = 0):
    """Droplet labels (src/cluster.cpp:17-82): int32 array of the field's
    (Nx, Ny, Nz, nC) shape, Fortran order, and the number of droplets."""
    a = _f(a)
    if a.ndim == 3:
        a = a[..., None]
    nx, ny, nz, nc = a.shape
    labels = np.zeros(a.shape, dtype=np.int32, order="F")
    n = lib().or_decompose(_ptr(a), nx, ny, nz, nc, int(component), float(cutoff), _ptr(labels))
    if n < 0:
        raise MemoryError("or_decompose")
    return labels, n


def uniform(seed: int, count: int, start: int = 0) -> np.ndarray:
    """splitmix64(seed*2^40 + i) mapped to [-1, 1) (same stream as gp5_amd.synth)."""
    out = np.empty(int(count), dtype=np.float64)
    lib().or_uniform(int(seed), int(start), int(count), _ptr(out))
    return out


def momentum(i: int, n: int, dk: float) -> float:
    return lib().or_momentum(int(i), int(n), float(dk))
