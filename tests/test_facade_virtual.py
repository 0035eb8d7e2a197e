"""User-defined functionals and constraints in the C++ operator API.

The reference's steppers call the functional and the constraint through virtual
methods (src/functional.h:15, src/constraint.h:11, src/stepper.cpp:30-70), so a
class derived by the user changes what a step computes.  gp::stepper takes the
fused device path only for the built-in kinds and otherwise runs the step stage
by stage through the user's apply(); tests/cpp/custom_functional.cpp defines an
LHY functional with a constant energy shift and its own phase constraint, whose
runs must equal the oracle's LHY runs with a uniform potential and the phase
constraint."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT, rel_l2

SHAPE, LEFT, RIGHT = (24, 20, 16), (-5.0, -5.0, -4.0), (5.0, 5.0, 4.0)


def _build(out_dir):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = os.path.join(str(out_dir), "custom_functional")
    cmd = [gxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "gp5_amd", "csrc"),
           os.path.join(ROOT, "tests", "cpp", "custom_functional.cpp"), "-L", os.path.join(ROOT, "gp5_amd"), "-lgp5",
           "-Wl,-rpath," + os.path.join(ROOT, "gp5_amd"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


def test_custom_functional_example_compiles(tmp_path):
    assert os.path.exists(_build(tmp_path))


def _initial():
    from gp5_amd import synth

    x, y, z = synth.axes(SHAPE, LEFT, RIGHT)
    r2 = x[:, None, None] ** 2 + y[None, :, None] ** 2 + z[None, None, :] ** 2
    a = np.exp(-r2 / 4.0) * (1.0 + 0.1j * np.sin(x)[:, None, None])
    phi = np.arctan2(y[None, :, None] + 0 * x[:, None, None] + 0 * z[None, None, :], x[:, None, None] + 0.01)
    return np.asfortranarray(a[..., None]), np.asfortranarray(phi[..., None])


@pytest.mark.gpu
@pytest.mark.parametrize("case", [0, 1, 2])
def test_user_functional_and_constraint_are_honoured(tmp_path, oracle_mod, case):
    exe = _build(tmp_path)
    out = os.path.join(str(tmp_path), "psi.bin")
    r = subprocess.run([exe, str(case), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    fused, fcalls, ccalls = int(words[1]), int(words[3]), int(words[5])
    got = np.fromfile(out, dtype=np.complex128).reshape(SHAPE + (1,), order="F")
    O = oracle_mod
    a0, phi = _initial()
    geo = O.Geometry(SHAPE, LEFT, RIGHT, 1)
    V = np.full(SHAPE + (1,), 0.35, order="F")
    if case == 0:
        assert (fused, fcalls, ccalls) == (0, 4 * 6, 6)
        ref = O.advance(O.Functional(kind=O.OR_LHY, V=V), a0, geo, stepper="RK4", dt=1e-3, imaginary=True, nsteps=6,
                        renormalize=[2.0], phase=phi)
    elif case == 1:
        assert (fused, fcalls, ccalls) == (0, 8, 0)
        ref = O.advance(O.Functional(kind=O.OR_LHY, V=V), a0, geo, stepper="eulero", dt=1e-3, imaginary=False,
                        nsteps=8)
    else:
        assert (fused, fcalls) == (1, 0)
        ref = O.advance(O.Functional(kind=O.OR_LHY), a0, geo, stepper="RK4", dt=1e-3, imaginary=False, nsteps=5)
    assert rel_l2(got, ref) <= 1e-12
