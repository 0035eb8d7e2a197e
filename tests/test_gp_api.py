"""The Python API with the reference's names (gp5_amd.gp over the gpCpp module), tested
the way the reference tests its own Python layer (python/test/test.py:9-221), plus
parity of the API's apply/advance with the oracle.

CPU tests need no GPU (geometry, module surface, errors); the rest run the HIP path."""
import os

import numpy as np
import pytest

from conftest import rel_l2, seeded_field

gp = pytest.importorskip("gp5_amd.gp")


# ---------------------------------------------------------------- CPU ----------

def test_module_surface():
    from gp5_amd import gpCpp

    for name in ["geometry", "field", "model", "LHY", "GP", "UnlockedLHY", "timeStepper", "decomposition"]:
        assert hasattr(gpCpp, name), name
    for meth in ["getLocalShape", "getTensor", "setTensor", "getOffset", "getNorm", "normalize", "save", "load"]:
        assert hasattr(gpCpp.field, meth), meth
    for meth in ["setRenormalization", "unsetRenormalization", "setFunctional", "advance", "getTime"]:
        assert hasattr(gpCpp.timeStepper, meth), meth
    assert hasattr(gpCpp.model, "apply")
    assert hasattr(gpCpp.decomposition, "decompose")


def test_geo():
    """python/test/test.py:11-18"""
    geo = gp.geometry((100, 200, 50))
    assert geo.shape == [100, 200, 50]
    assert geo.left == [-1, -1, -1]
    assert geo.right == [1, 1, 1]
    assert geo.spaceStep[0] == 2 / 100
    assert geo.spaceStep[1] == 2 / 200
    assert geo.spaceStep[2] == 2 / 50


def test_unknown_stepper_raises():
    from gp5_amd import gpCpp

    with pytest.raises(RuntimeError, match="Unknown stepper"):
        gpCpp.timeStepper(0.1, "RK2", True)
    s = gpCpp.timeStepper(0.1, "RK4", False)
    assert s.getTime() == 0.0


def test_processor_grid_must_be_single():
    geo = gp.geometry((8, 8, 8))
    with pytest.raises(ValueError, match="processorGrid"):
        gp.field(geo, processorGrid=[2, 1, 1])


def test_field_without_device_is_an_error_not_a_fallback():
    """With no GPU the field constructor raises; there is no CPU path behind the API."""
    geo = gp.geometry((8, 8, 8))
    try:
        psi = gp.field(geo)
    except RuntimeError as e:
        assert "device" in str(e) or "HIP" in str(e)
    else:  # a GPU is present: the field lives in HBM
        assert psi.localShape == [8, 8, 8]


# ---------------------------------------------------------------- GPU ----------

@pytest.mark.gpu
def test_field():
    """python/test/test.py:21-46 (one process)"""
    geo = gp.geometry((100, 200, 50))
    psi = gp.field(geo, nComponents=1, processorGrid=[1, 1, 1])
    assert psi.localShape == [100, 200, 50]
    for d in range(3):
        assert psi.array.shape[d] == psi.localShape[d]
    X, Y, Z = psi.grid()
    assert np.min(X) == pytest.approx(geo.left[0] + geo.spaceStep[0] / 2)
    assert np.max(X) == pytest.approx(geo.right[0] - geo.spaceStep[0] / 2)
    assert X[1, 0, 0, 0] - X[0, 0, 0, 0] == pytest.approx(geo.spaceStep[0])
    psi.array = X**2 + Y**2 + Z**2
    assert np.sum(psi.array - (X**2 + Y**2 + Z**2)) == pytest.approx(0)
    assert np.array_equal(psi.array, (X**2 + Y**2 + Z**2).astype(np.complex128))


@pytest.mark.gpu
def test_set_tensor_rank_error():
    psi = gp.field(gp.geometry((8, 8, 8)))
    with pytest.raises(RuntimeError, match="Unexpected rank"):
        psi._fieldCpp.setTensor(np.zeros((8, 8), dtype=np.complex128))


@pytest.mark.gpu
def test_LHY():
    """python/test/test.py:73-89"""
    psi = gp.field(gp.geometry((100, 200, 50)), nComponents=1, processorGrid=[1, 1, 1])
    model = gp.LHY(psi)
    assert model.nComponents == 1


@pytest.mark.gpu
def test_LHY_apply():
    """python/test/test.py:92-119: H exp(-r^2) in closed form, max error < 1e-4."""
    geo = gp.geometry((100, 100, 100), left=[-4, -4, -4], right=[4, 4, 4])
    psi1 = gp.field(geo, nComponents=1)
    psi2 = gp.field(geo, nComponents=1)
    model = gp.LHY(psi1)
    X, Y, Z = psi1.grid()
    r = np.sqrt(X**2 + Y**2 + Z**2)
    assert model.nComponents == 1
    psi1.array = np.exp(-r**2) + 1j * 0
    model.apply(psi1, psi2, t=0)
    alpha = 1
    p = np.real(psi1.array)
    psi_expected = alpha * (3 - 2 * alpha * r**2) * p + (-3 * p**2 + 5 / 2 * p**3) * p
    assert np.max(np.abs(psi_expected - np.real(psi2.array))) < 1e-4


@pytest.mark.gpu
def test_renormalize():
    """python/test/test.py:123-147"""
    geo = gp.geometry((50, 50, 50), left=[-5, -5, -5], right=[5, 5, 5])
    psi = gp.field(geo, nComponents=1)
    X, Y, Z = psi.grid()
    r = np.sqrt(X**2 + Y**2 + Z**2)
    psi0 = np.exp(-r**2)
    psi.array = psi0
    psi.normalize(1, component=0)
    assert psi.norm(0) == pytest.approx(1, abs=1e-7)
    psi.array = psi0
    psi.normalize([1])
    N = psi.norm()
    assert len(N) == 1
    assert N[0] == pytest.approx(1, abs=1e-7)
    # a resolved Gaussian: sum |psi0|^2 dV = (pi/2)^{3/2}
    psi.array = psi0
    assert psi.norm(0) == pytest.approx((np.pi / 2) ** 1.5, rel=1e-12)


@pytest.mark.gpu
def test_stepper(oracle_mod):
    """python/test/test.py:149-197 (400 RK4 imaginary-time steps with normalize before each);
    the reference prints mu per block; here each block's mu is checked against the oracle."""
    O = oracle_mod
    geo = gp.geometry((50, 50, 50), left=[-32, -32, -32], right=[32, 32, 32])
    psi = gp.field(geo, nComponents=1)
    X, Y, Z = psi.grid()
    r = np.sqrt(X**2 + Y**2 + Z**2)
    psi.array = np.exp(-r**2 / (2 * 5**2))
    N = 370
    psi.normalize([N])
    model = gp.LHY(psi)
    stepper = gp.stepper(timeStep=0.1 * geo.spaceStep[0] ** 2, name="RK4", model=model)

    ogeo = O.Geometry((50, 50, 50), (-32, -32, -32), (32, 32, 32), 1)
    F = O.Functional(kind=O.OR_LHY)
    ref = psi.array
    maxDensityOp = gp.maxDensity()
    nBlocks, stepsPerBlock = 2, 50
    for _ in range(nBlocks):
        for _ in range(stepsPerBlock):
            psi.normalize([N])
            stepper.advance(psi, nSteps=1)
            ref = O.normalize(ref, ogeo, 0, N)
            ref = O.advance(F, ref, ogeo, stepper="RK4", dt=stepper.timeStep, imaginary=True, nsteps=1)
        mu = 0.5 * np.log(N / psi.norm(0)) / stepper.timeStep
        mu_ref = 0.5 * np.log(N / O.norm(ref, ogeo, 0)) / stepper.timeStep
        assert abs(mu - mu_ref) < 1e-9
        psi.normalize([N])
        ref = O.normalize(ref, ogeo, 0, N)
        assert maxDensityOp(psi, root=0) == pytest.approx(np.max(np.abs(ref)) ** 2, rel=1e-12)
    assert rel_l2(psi.array, ref) < 1e-12
    assert stepper.time == pytest.approx(nBlocks * stepsPerBlock * stepper.timeStep, rel=1e-12)


@pytest.mark.gpu
def test_io(tmp_path):
    """python/test/test.py:200-217, with numpy files instead of parallel HDF5."""
    geo = gp.geometry((50, 50, 50), left=[-32, -32, -32], right=[32, 32, 32])
    psi = gp.field(geo, nComponents=1)
    X, Y, Z = psi.grid()
    r = np.sqrt(X**2 + Y**2 + Z**2)
    psi0 = np.exp(-r**2 / (2 * 5**2))
    psi.array = psi0
    fn = os.path.join(str(tmp_path), "out", "test.npz")
    psi.save(fn)
    psi.array = psi0 * 0
    psi.load(fn)
    assert np.sum(np.abs(psi.array - psi0)) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("name,imaginary", [("RK4", False), ("eulero", True), ("Eulero", False)])
def test_stepper_parity(oracle_mod, name, imaginary):
    """timeStepper.advance(field, n) against the oracle, seeded droplet-like input, 24x20x16."""
    O = oracle_mod
    shape, left, right = (24, 20, 16), (-6.0, -5.0, -4.0), (6.0, 5.0, 4.0)
    geo = gp.geometry(shape, left=list(left), right=list(right))
    psi = gp.field(geo)
    a0 = seeded_field(shape, 1, 11, 0.5)
    psi.array = a0
    model = gp.LHY(psi)
    dt = 2e-3
    st = gp.stepper(model, dt, name=name, imaginary=imaginary)
    st.advance(psi, nSteps=7)
    ref = O.advance(O.Functional(kind=O.OR_LHY), a0, O.Geometry(shape, left, right, 1),
                    stepper="eulero" if name != "RK4" else "RK4", dt=dt, imaginary=imaginary, nsteps=7)
    assert rel_l2(psi.array, ref) < 1e-12
    assert st.time == pytest.approx(7 * dt)


@pytest.mark.gpu
def test_renormalized_stepper_parity(oracle_mod):
    """setRenormalization([N]): per-stage normalisation (src/stepper.cpp:7-22,46-70)."""
    O = oracle_mod
    shape, left, right = (16, 16, 16), (-5.0, -5.0, -5.0), (5.0, 5.0, 5.0)
    geo = gp.geometry(shape, left=list(left), right=list(right))
    psi = gp.field(geo)
    a0 = seeded_field(shape, 1, 12, 0.3)
    psi.array = a0
    st = gp.stepper(gp.LHY(psi), 3e-3, name="RK4", imaginary=True)
    st.setRenormalization([5.0])
    st.advance(psi, nSteps=5)
    ref = O.advance(O.Functional(kind=O.OR_LHY), a0, O.Geometry(shape, left, right, 1), stepper="RK4",
                    dt=3e-3, imaginary=True, nsteps=5, renormalize=[5.0])
    assert rel_l2(psi.array, ref) < 1e-12
    assert psi.norm(0) == pytest.approx(5.0, rel=1e-12)


@pytest.mark.gpu
def test_gp_and_unlocked_models_apply(oracle_mod):
    O = oracle_mod
    shape, left, right = (20, 18, 16), (-4.0, -4.0, -4.0), (4.0, 4.0, 4.0)
    geo = gp.geometry(shape, left=list(left), right=list(right))
    ogeo = O.Geometry(shape, left, right, 2)
    a0 = seeded_field(shape, 2, 13, 0.4)
    src, dst = gp.field(geo, nComponents=2), gp.field(geo, nComponents=2)
    src.array = a0

    g = [[1.0, 0.3], [0.3, 0.8]]
    m = gp.GP(src, couplings=g, masses=[1.0, 1.5], omegas=[[1.0, 1.2, 0.8], [0.5, 0.5, 0.5]])
    m.apply(src, dst)
    X, Y, Z = src.grid()
    V = np.empty(shape + (2,))
    for c, w in enumerate([[1.0, 1.2, 0.8], [0.5, 0.5, 0.5]]):
        V[..., c] = 0.5 * (w[0] ** 2 * X[..., c] ** 2 + w[1] ** 2 * Y[..., c] ** 2 + w[2] ** 2 * Z[..., c] ** 2)
    ref = O.apply(O.Functional(kind=O.OR_GP, couplings=np.array(g), masses=(1.0, 1.5), V=V), a0, ogeo)
    assert rel_l2(dst.array, ref) < 1e-12
    assert m.nComponents == 2

    u = gp.UnlockedLHY(src, 0.6, 0.9, -1.1)
    u.apply(src, dst)
    ref = O.apply(O.Functional(kind=O.OR_UNLOCKED, alpha=0.6, beta=0.9, eta=-1.1), a0, ogeo)
    assert rel_l2(dst.array, ref) < 1e-12
    # the unlocked functional is not a derivative of an energy unless beta = 1: no energy
    with pytest.raises(RuntimeError, match="no energy"):
        u.energy(src)


@pytest.mark.gpu
def test_advance_log_matches_oracle_norms(oracle_mod):
    O = oracle_mod
    shape, left, right = (16, 16, 16), (-8.0, -8.0, -8.0), (8.0, 8.0, 8.0)
    geo = gp.geometry(shape, left=list(left), right=list(right))
    psi = gp.field(geo)
    a0 = seeded_field(shape, 1, 14, 0.5)
    psi.array = a0
    model = gp.LHY(psi)
    st = gp.stepper(model, 1e-3, name="RK4", imaginary=False)
    norms, ens = st.advance_log(psi, 3)
    ogeo = O.Geometry(shape, left, right, 1)
    F = O.Functional(kind=O.OR_LHY)
    a = a0
    for i in range(3):
        assert norms[i, 0] == pytest.approx(O.norm(a, ogeo, 0), rel=1e-12)
        assert ens[i] == pytest.approx(O.energy(F, a, ogeo), rel=1e-11)
        a = O.advance(F, a, ogeo, stepper="RK4", dt=1e-3, imaginary=False, nsteps=1)
    assert rel_l2(psi.array, a) < 1e-12
