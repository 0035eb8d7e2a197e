"""Python API with the names of the reference's `gp` package
(python/gp/gp.py:11-222): geometry, field, LHY, stepper, maxDensity, on top
of the gpCpp extension (HIP kernels on one MI355X).

    from gp5_amd import gp
    geo = gp.geometry((30, 30, 30), left=[-10] * 3, right=[10] * 3)
    psi = gp.field(geo)
    ...

Differences from the reference, all on the host side:
  * mpi4py and netCDF4 are optional (one GPU, one process);
  * field.save / field.load read and write the reference's HDF5 field file
    serially (one GPU) and also .npy/.npz; the netCDF4 writer is not part of
    this build;
  * stepper accepts "RK4" and "eulero" (also spelled "Eulero", the name the
    reference documents); an unknown name raises;
  * additions: GP / UnlockedLHY models, model.energy, field.maxDensity,
    stepper.advance_log.
"""
from __future__ import annotations

import os
import re
from pathlib import Path

import numpy as np

from . import capi, gpCpp


class geometry:
    def __init__(self, shape, left=(-1, -1, -1), right=(1, 1, 1)):
        self._geometryCpp = gpCpp.geometry(list(left), list(right), list(shape))

    @property
    def left(self):
        return self._geometryCpp.getLeft()

    @property
    def right(self):
        return self._geometryCpp.getRight()

    @property
    def shape(self):
        return self._geometryCpp.getShape()

    @property
    def spaceStep(self):
        return [(self.right[d] - self.left[d]) / self.shape[d] for d in range(len(self.shape))]


class field:
    def __init__(self, geo, nComponents=1, processorGrid=(1, 1, 1), fftOrdering=(2, 0, 1)):
        if list(processorGrid) != [1, 1, 1]:
            raise ValueError("one GPU: processorGrid must be [1, 1, 1]")
        self._nComponents = nComponents
        self._geometry = geo
        self._fieldCpp = gpCpp.field(geo._geometryCpp, nComponents, list(processorGrid), list(fftOrdering))

    @property
    def localShape(self):
        return self._fieldCpp.getLocalShape()

    @property
    def array(self):
        return self._fieldCpp.getTensor()

    @array.setter
    def array(self, x):
        x = np.asarray(x)
        if x.ndim == 3 and self._nComponents == 1:
            x = x[..., None]
        self._fieldCpp.setTensor(np.asarray(x, dtype=np.complex128))

    @property
    def nComponents(self):
        return self._nComponents

    def norm(self, component=None):
        if component is not None:
            return self._fieldCpp.getNorm(component)
        return [self.norm(c) for c in range(self.nComponents)]

    def normalize(self, N, component=None):
        if component is not None:
            self._fieldCpp.normalize(component, N)
        else:
            for Nc, c in zip(N, range(self.nComponents)):
                self.normalize(Nc, component=c)

    def maxDensity(self, component=0):
        return self._fieldCpp.maxDensity(component)

    @property
    def offset(self):
        return self._fieldCpp.getOffset()

    def grid(self):
        """Cell-centred coordinates (X, Y, Z), each of shape (Nx, Ny, Nz, nC)."""
        meshShape = self.localShape
        left = self._geometry.left
        deltax = self._geometry.spaceStep
        axes = [(np.arange(0, meshShape[d]) + 0.5) * deltax[d] + left[d] + self.offset[d] * deltax[d]
                for d in range(3)]
        axes.append([1.0 for _ in range(self._nComponents)])
        X, Y, Z, _ = np.meshgrid(*axes, indexing="ij")
        return X, Y, Z

    def save(self, filename, format=None):
        """Format from the extension as python/gp/gp.py:98-117 guesses it: .hdf5 / .h5
        write the reference's HDF5 field file (src/io.cpp:56-152); .npz / .npy are
        additions; netCDF4 (.nc) is not part of this build."""
        if format is None:
            for pattern, fmt in ((r".*\.nc$", "netCDF4"), (r".*\.hdf5$", "hdf5"), (r".*\.h5$", "hdf5"),
                                 (r".*\.npz$", "npz"), (r".*\.npy$", "npy")):
                if re.match(pattern, filename):
                    format = fmt
                    break
        if format is None:
            raise RuntimeError(f"Could not guess format of {filename}")
        Path(os.path.dirname(filename) or ".").mkdir(parents=True, exist_ok=True)
        if format == "hdf5":
            capi.h5_write(filename, self.array)
        elif format == "npz":
            np.savez(filename, field=self.array)
        elif format == "npy":
            np.save(filename, self.array)
        elif format == "netCDF4":
            raise RuntimeError("netCDF4 output is not part of this build (use .hdf5)")
        else:
            raise RuntimeError(f"Unkown format {format}")

    def load(self, filename):
        """HDF5 field file (src/io.cpp:154-254) unless the name ends in .npz / .npy."""
        if filename.endswith(".npz"):
            with np.load(filename, allow_pickle=False) as z:
                self.array = z["field"]
        elif filename.endswith(".npy"):
            self.array = np.load(filename, allow_pickle=False)
        else:
            shape = tuple(self.localShape)
            self.array = capi.h5_read(filename, shape, self._nComponents)


class _model:
    def apply(self, psi1, psi2, t=0.0):
        self._funcCpp.apply(psi1._fieldCpp, psi2._fieldCpp, t)

    def energy(self, psi):
        return self._funcCpp.energy(psi._fieldCpp)

    @property
    def nComponents(self):
        return self._funcCpp.nComponents()


class LHY(_model):
    """LHYDropletFunctional: -1/2 lap + (a rho + b rho^{3/2}), a = -3, b = 5/2."""

    def __init__(self, psi, a=-3.0, b=2.5):
        if psi.nComponents != 1:
            raise ValueError("Field should have one component")
        self._funcCpp = gpCpp.LHY(psi._fieldCpp, a, b)


class GP(_model):
    """gpFunctional: -1/2 m^-1 lap + V + contact couplings (src/functional.cpp:13-102)."""

    def __init__(self, psi, couplings=None, masses=None, omegas=None):
        self._funcCpp = gpCpp.GP(psi._fieldCpp, [] if couplings is None else [list(r) for r in couplings],
                                 [] if masses is None else list(masses),
                                 [] if omegas is None else [list(w) for w in omegas])


class UnlockedLHY(_model):
    """LHYDropletUnlockedFunctional (src/functional.cpp:125-150), two components."""

    def __init__(self, psi, alpha, beta, eta):
        self._funcCpp = gpCpp.UnlockedLHY(psi._fieldCpp, alpha, beta, eta)


class stepper:
    def __init__(self, model, timeStep, name="RK4", imaginary=True, renormalize=False):
        """name: str, allowed: RK4, eulero"""
        self._timeStep = timeStep
        self._name = "eulero" if name in ("eulero", "Eulero") else name
        self._imaginary = imaginary
        self._stepperCpp = gpCpp.timeStepper(timeStep, self._name, imaginary)
        self._model = model
        self._stepperCpp.setFunctional(model._funcCpp)
        self._renormalize = False
        self._stepperCpp.unsetRenormalization()

    def setRenormalization(self, N):
        self._stepperCpp.setRenormalization(list(N))

    def advance(self, psi, nSteps=1):
        self._stepperCpp.advance(psi._fieldCpp, nSteps)

    def advance_log(self, psi, nSteps=1):
        """advance, returning (norms[nSteps, nC], energies[nSteps]) at the start of each step."""
        return self._stepperCpp.advanceLog(psi._fieldCpp, nSteps)

    @property
    def time(self):
        return self._stepperCpp.getTime()

    @property
    def timeStep(self):
        return self._timeStep


class maxDensity:
    """max |psi|^2 over the field (python/gp/gp.py:210-222; one process, no MPI reduce)."""

    def __call__(self, psi, root=0):
        return max(psi.maxDensity(c) for c in range(psi.nComponents))
