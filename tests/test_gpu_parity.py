"""HIP path vs the CPU oracle, through the C ABI (libgp5.so via ctypes).

Bar: relative L2 <= 1e-12 for stepped trajectories (floating point; FP64 RK4
at 256^3 must match the reference to 1e-12), <= 1e-13 for single transforms.  The
reference's own known-answer tests are reproduced with their tolerances and
with the tighter FP64 bar.
"""
import numpy as np
import pytest

from conftest import rel_l2, seeded_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def C():
    from gp5_amd import capi

    capi.load()
    return capi


def _ctx(C, shape, left, right, nc=1):
    return C.Context(shape, left, right, nc)


FFT_CASES = [((16, 16, 16), 1), ((8, 12, 10), 2), ((30, 30, 30), 1), ((7, 11, 17), 1), ((64, 1, 5), 1),
             ((96, 24, 40), 1), ((256, 16, 8), 1), ((512, 4, 8), 2), ((32, 128, 64), 1), ((20, 64, 512), 1),
             ((16, 256, 32), 2), ((128, 3, 256), 1), ((192, 12, 8), 1), ((16, 96, 20), 2), ((12, 10, 192), 1),
             ((96, 192, 4), 1)]


@pytest.mark.parametrize("shape,nc", FFT_CASES)
def test_fft_matches_oracle(C, oracle_mod, shape, nc):
    print(f"fft {shape} nc={nc}", flush=True)
    a = seeded_field(shape, nc, seed=11)
    ctx = _ctx(C, shape, (-1, -1, -1), (1, 1, 1), nc)
    src, dst = ctx.field(a), ctx.field()
    for direction, sign in ((C.GP5_FORWARD, -1), (C.GP5_BACKWARD, +1)):
        C.fft(ctx, src, dst, direction)
        ref = oracle_mod.fft3d(a, sign)
        assert rel_l2(dst.download_complex(), ref) <= 1e-13
    # input preserved (out-of-place semantics)
    np.testing.assert_array_equal(src.download_complex(), a)


def test_fft_forward_backward_reference(C):
    """tst/fft_test.cpp:10-83: Gaussian sigma=(1,0.5,0.2) on 50x170x150 over [-1/2,1/2]^3."""
    shape, left, right = (50, 170, 150), (-0.5,) * 3, (0.5,) * 3
    from gp5_amd import synth

    x, y, z = synth.axes(shape, left, right)
    g = np.exp(-(x[:, None, None] ** 2 / (2 * 1.0 ** 2) + y[None, :, None] ** 2 / (2 * 0.5 ** 2)
                 + z[None, None, :] ** 2 / (2 * 0.2 ** 2)))
    a = np.asfortranarray(g[..., None].astype(np.complex128))
    ctx = _ctx(C, shape, left, right)
    f1, f2, f3 = ctx.field(a), ctx.field(), ctx.field()
    C.fft(ctx, f1, f2, C.GP5_FORWARD)
    C.fft(ctx, f2, f3, C.GP5_BACKWARD)
    back = f3.download_complex() / np.prod(shape)
    err = np.max(np.abs(back - a))
    assert err <= 1e-6  # reference TOL (src/traits.h:46)
    assert err <= 1e-14


def test_laplacian_closed_form_reference(C):
    """tst/fft_test.cpp:85-153: spectral Laplacian of a sigma=0.05 Gaussian on 300x250x77."""
    shape, left, right = (300, 250, 77), (-0.5,) * 3, (0.5,) * 3
    from gp5_amd import synth

    sigma = 0.05
    alpha = 1 / (2 * sigma * sigma)
    r2 = synth.r2_grid(shape, left, right)
    a = np.asfortranarray(np.exp(-alpha * r2)[..., None].astype(np.complex128))
    expect = np.exp(-alpha * r2) * (-r2 * alpha * 2 + 3) * (-2 * alpha)
    ctx = _ctx(C, shape, left, right)
    f, L = ctx.field(a), ctx.field()
    C.laplacian(ctx, f, L)
    err = np.max(np.abs(L.download_complex()[..., 0] - expect))
    assert err <= 1e-6  # reference TOL
    assert err <= 1e-9 * np.max(np.abs(expect))


@pytest.mark.parametrize("shape,nc", [((16, 16, 16), 1), ((12, 10, 9), 2), ((32, 24, 20), 1)])
def test_laplacian_matches_oracle(C, oracle_mod, shape, nc):
    left, right = (-3.0, -2.0, -4.0), (3.0, 2.5, 4.0)
    a = seeded_field(shape, nc, seed=5)
    geo = oracle_mod.Geometry(shape, left, right, nc)
    ref = oracle_mod.laplacian(a, geo)
    ctx = _ctx(C, shape, left, right, nc)
    f, L = ctx.field(a), ctx.field()
    C.laplacian(ctx, f, L)
    assert rel_l2(L.download_complex(), ref) <= 1e-13


def _smooth(shape, left, right, nc, seed, amp=1.0):
    from gp5_amd import synth

    r2 = synth.r2_grid(shape, left, right)
    out = np.empty(tuple(shape) + (nc,), dtype=np.complex128, order="F")
    for c in range(nc):
        noise = seeded_field(shape, 1, seed + c)[..., 0]
        out[..., c] = amp * np.exp(-r2 / (2 * (1.5 + 0.3 * c) ** 2)) * (1 + 0.05 * noise)
    return out


def _functional_cases(oracle_mod, shape, left, right):
    from gp5_amd import synth

    V1 = synth.harmonic_potential(shape, left, right, [(1.0, 0.8, 1.2)])
    V2 = synth.harmonic_potential(shape, left, right, [(1.0, 1.0, 1.0), (0.5, 0.7, 0.9)])
    return {
        "lhy": (1, dict(kind=1), dict(kind=oracle_mod.OR_LHY)),
        "lhy_trap": (1, dict(kind=1, V=V1), dict(kind=oracle_mod.OR_LHY, V=V1)),
        "gp_g": (1, dict(kind=0, couplings=[[1.0]], V=V1),
                 dict(kind=oracle_mod.OR_GP, couplings=np.array([[1.0]]), V=V1)),
        "gp_free": (1, dict(kind=0), dict(kind=oracle_mod.OR_GP)),
        "gp_2c": (2, dict(kind=0, couplings=[[1.0, 0.4], [0.4, 0.7]], masses=(1.0, 2.0), V=V2),
                  dict(kind=oracle_mod.OR_GP, couplings=np.array([[1.0, 0.4], [0.4, 0.7]]), masses=(1.0, 2.0), V=V2)),
        "unlocked": (2, dict(kind=2, alpha=0.6, beta=0.8, eta=-0.9),
                     dict(kind=oracle_mod.OR_UNLOCKED, alpha=0.6, beta=0.8, eta=-0.9)),
    }


def _desc(C, ctx, spec):
    spec = dict(spec)
    V = spec.pop("V", None)
    vb = ctx.real(V) if V is not None else None
    return C.functional_desc(potential=vb, **spec), vb


@pytest.mark.parametrize("case", ["lhy", "lhy_trap", "gp_g", "gp_free", "gp_2c", "unlocked"])
def test_functional_apply_matches_oracle(C, oracle_mod, case):
    shape, left, right = (24, 20, 16), (-5.0, -5.0, -4.0), (5.0, 5.0, 4.0)
    nc, gspec, ospec = _functional_cases(oracle_mod, shape, left, right)[case]
    a = _smooth(shape, left, right, nc, seed=21)
    geo = oracle_mod.Geometry(shape, left, right, nc)
    ref = oracle_mod.apply(oracle_mod.Functional(**ospec), a, geo)
    ctx = _ctx(C, shape, left, right, nc)
    f, vb = _desc(C, ctx, gspec)
    src, dst = ctx.field(a), ctx.field()
    C.functional_apply(ctx, f, src, dst, 0.0)
    assert rel_l2(dst.download_complex(), ref) <= 1e-13


def test_lhy_apply_closed_form_reference(C):
    """python/test/test.py:92-119: psi = exp(-r^2) on 100^3 over [-4,4]^3."""
    shape, left, right = (100, 100, 100), (-4.0,) * 3, (4.0,) * 3
    from gp5_amd import synth

    r2 = synth.r2_grid(shape, left, right)
    psi = np.exp(-r2)
    ctx = _ctx(C, shape, left, right)
    src, dst = ctx.field(psi[..., None].astype(np.complex128)), ctx.field()
    C.functional_apply(ctx, C.functional_desc(C.GP5_FUNC_LHY), src, dst, 0.0)
    expect = (3 - 2 * r2) * psi + (-3 * psi ** 2 + 2.5 * psi ** 3) * psi
    out = dst.download_complex()[..., 0]
    assert np.max(np.abs(expect - out.real)) < 1e-4
    assert np.max(np.abs(out.imag)) < 1e-12


def test_norm_normalize_reference(C):
    """python/test/test.py:123-147, plus the closed form int e^{-2 r^2} = (pi/2)^{3/2}."""
    shape, left, right = (50, 50, 50), (-5.0,) * 3, (5.0,) * 3
    from gp5_amd import synth

    psi = np.exp(-synth.r2_grid(shape, left, right))[..., None].astype(np.complex128)
    ctx = _ctx(C, shape, left, right)
    f = ctx.field(psi)
    assert abs(C.norm(ctx, f, 0) - (np.pi / 2) ** 1.5) < 1e-13
    C.normalize(ctx, f, 0, 1.0)
    assert abs(C.norm(ctx, f, 0) - 1.0) < 1e-14
    assert abs(C.max_density(ctx, f, 0) - np.max(np.abs(f.download_complex()) ** 2)) < 1e-15


STEP_CASES = [
    ("eulero", True, None, "lhy"), ("eulero", False, None, "gp_g"), ("eulero", True, [3.0], "gp_g"),
    ("RK4", True, None, "lhy"), ("RK4", False, None, "lhy"), ("RK4", True, [2.0], "lhy_trap"),
    ("RK4", False, None, "gp_2c"), ("RK4", True, [1.5, 0.5], "gp_2c"), ("RK4", False, None, "unlocked"),
    ("eulero", True, [1.0, 2.0], "unlocked"),
]


@pytest.mark.parametrize("shape", [(24, 20, 16), (32, 16, 20), (96, 12, 24)], ids=["generic_x", "fused_x", "radix3_x"])
@pytest.mark.parametrize("name,imag,renorm,case", STEP_CASES)
def test_advance_matches_oracle(C, oracle_mod, name, imag, renorm, case, shape):
    print(f"advance {name} imag={imag} renorm={renorm} {case} {shape}", flush=True)
    left, right = (-5.0, -5.0, -4.0), (5.0, 5.0, 4.0)
    nc, gspec, ospec = _functional_cases(oracle_mod, shape, left, right)[case]
    a = _smooth(shape, left, right, nc, seed=31)
    dt, nsteps = 2e-3, 12
    geo = oracle_mod.Geometry(shape, left, right, nc)
    ref = oracle_mod.advance(oracle_mod.Functional(**ospec), a, geo, stepper=name, dt=dt, imaginary=imag,
                             nsteps=nsteps, renormalize=renorm)
    ctx = _ctx(C, shape, left, right, nc)
    f, vb = _desc(C, ctx, gspec)
    psi = ctx.field(a)
    kind = C.GP5_STEP_RK4 if name == "RK4" else C.GP5_STEP_EULERO
    s = C.stepper_desc(kind, dt=dt, imaginary=imag, renormalize=renorm)
    C.advance(ctx, f, s, psi, nsteps)
    assert rel_l2(psi.download_complex(), ref) <= 1e-12


@pytest.mark.parametrize("shape", [(256, 8, 12), (8, 256, 12), (12, 8, 256)], ids=["x256", "y256", "z256"])
@pytest.mark.parametrize("name,imag,renorm,case", STEP_CASES)
def test_advance_256_axes_match_oracle(C, oracle_mod, name, imag, renorm, case, shape):
    """The 256-point register passes of the 256^3 workload in every stage mode: the fused
    x epilogue (component-local and coupled variants, all four RK stages, Euler, with
    per-stage renormalisation), the 256-point y pass and the fused z multiplier pass.
    dt = 2e-4 keeps dt k_max^2 / 2 = 0.65 inside the explicit limit (256 points over 10)."""
    print(f"advance256 {name} imag={imag} renorm={renorm} {case} {shape}", flush=True)
    left, right = (-5.0, -5.0, -4.0), (5.0, 5.0, 4.0)
    nc, gspec, ospec = _functional_cases(oracle_mod, shape, left, right)[case]
    a = _smooth(shape, left, right, nc, seed=37)
    dt, nsteps = 2e-4, 12
    geo = oracle_mod.Geometry(shape, left, right, nc)
    ref = oracle_mod.advance(oracle_mod.Functional(**ospec), a, geo, stepper=name, dt=dt, imaginary=imag,
                             nsteps=nsteps, renormalize=renorm)
    ctx = _ctx(C, shape, left, right, nc)
    f, vb = _desc(C, ctx, gspec)
    psi = ctx.field(a)
    kind = C.GP5_STEP_RK4 if name == "RK4" else C.GP5_STEP_EULERO
    s = C.stepper_desc(kind, dt=dt, imaginary=imag, renormalize=renorm)
    C.advance(ctx, f, s, psi, nsteps)
    assert rel_l2(psi.download_complex(), ref) <= 1e-12


@pytest.mark.parametrize("shape", [(16, 12, 192), (192, 8, 12), (12, 96, 96)])
def test_radix3_register_passes_rk4(C, oracle_mod, shape):
    """RK4 real time through the register passes of lengths 96 and 192 on every axis role
    (fused x epilogue, y pass, fused z multiplier pass); dt below the explicit limit of the
    finest axis (k_max^2 / 2 ~ 2800 at 192 points over 8)."""
    left, right = (-4.0, -4.0, -4.0), (4.0, 4.0, 4.0)
    a = _smooth(shape, left, right, 1, seed=33)
    geo = oracle_mod.Geometry(shape, left, right, 1)
    F = oracle_mod.Functional(kind=oracle_mod.OR_LHY)
    ref = oracle_mod.advance(F, a, geo, stepper="RK4", dt=2e-4, imaginary=False, nsteps=10)
    ctx = _ctx(C, shape, left, right)
    psi = ctx.field(a)
    s = C.stepper_desc(C.GP5_STEP_RK4, dt=2e-4, imaginary=False)
    C.advance(ctx, C.functional_desc(C.GP5_FUNC_LHY), s, psi, 10)
    assert rel_l2(psi.download_complex(), ref) <= 1e-12


def test_phase_constraint_step_matches_oracle(C, oracle_mod):
    shape, left, right = (16, 16, 16), (-4.0,) * 3, (4.0,) * 3
    from gp5_amd import synth

    x, y, z = synth.axes(shape, left, right)
    phase = np.arctan2(y[None, :, None] + 0 * x[:, None, None] + 0 * z[None, None, :], x[:, None, None] + 0.01)
    phase = np.asfortranarray(phase[..., None])
    a = _smooth(shape, left, right, 1, seed=41)
    geo = oracle_mod.Geometry(shape, left, right, 1)
    ref = oracle_mod.advance(oracle_mod.Functional(kind=oracle_mod.OR_LHY), a, geo, stepper="RK4", dt=1e-3,
                             imaginary=True, nsteps=5, renormalize=[2.0], phase=phase)
    ctx = _ctx(C, shape, left, right)
    pb = ctx.real(phase)
    psi = ctx.field(a)
    s = C.stepper_desc(C.GP5_STEP_RK4, dt=1e-3, imaginary=True, renormalize=[2.0], phase=pb)
    C.advance(ctx, C.functional_desc(C.GP5_FUNC_LHY), s, psi, 5)
    assert rel_l2(psi.download_complex(), ref) <= 1e-12
    # standalone constraint
    b = ctx.field(a)
    C.phase_constraint(ctx, b, pb)
    expect = np.abs(a) * np.exp(1j * phase)
    assert rel_l2(b.download_complex(), expect) <= 1e-14


@pytest.mark.parametrize("case", ["lhy", "gp_g", "gp_2c"])
def test_energy_matches_oracle(C, oracle_mod, case):
    shape, left, right = (24, 20, 16), (-5.0, -5.0, -4.0), (5.0, 5.0, 4.0)
    nc, gspec, ospec = _functional_cases(oracle_mod, shape, left, right)[case]
    a = _smooth(shape, left, right, nc, seed=51)
    geo = oracle_mod.Geometry(shape, left, right, nc)
    ref = oracle_mod.energy(oracle_mod.Functional(**ospec), a, geo)
    ctx = _ctx(C, shape, left, right, nc)
    f, vb = _desc(C, ctx, gspec)
    e = C.energy(ctx, f, ctx.field(a))
    assert abs(e - ref) <= 1e-12 * max(1.0, abs(ref))


def test_config1_harmonic_euler_parity(C, oracle_mod):
    """64^3 on [-10,10]^3, gpFunctional g=1, harmonic omega=1, Euler imaginary dt=1e-3,
    renormalised to N=1 every step: 100 steps vs the oracle."""
    from gp5_amd import synth

    shape, left, right, psi0, V = synth.harmonic_64()
    geo = oracle_mod.Geometry(shape, left, right, 1)
    F = oracle_mod.Functional(kind=oracle_mod.OR_GP, couplings=np.array([[1.0]]), V=V)
    ref = oracle_mod.advance(F, psi0, geo, stepper="eulero", dt=1e-3, imaginary=True, nsteps=100, renormalize=[1.0])
    ctx = _ctx(C, shape, left, right)
    vb = ctx.real(V)
    f = C.functional_desc(C.GP5_FUNC_GP, couplings=[[1.0]], potential=vb)
    psi = ctx.field(psi0)
    s = C.stepper_desc(C.GP5_STEP_EULERO, dt=1e-3, imaginary=True, renormalize=[1.0])
    C.advance(ctx, f, s, psi, 100)
    assert rel_l2(psi.download_complex(), ref) <= 1e-12
    assert abs(C.norm(ctx, psi, 0) - 1.0) < 1e-13


def test_config2_droplet_rk4_parity_64(C, oracle_mod):
    """LHY droplet (-3, 2.5), box L=40, flat top r=6, N=3000, real-time RK4 dt=5e-4:
    100 steps at 64^3 vs the oracle."""
    from gp5_amd import synth

    shape, left, right, psi0 = synth.droplet_256(64)
    geo = oracle_mod.Geometry(shape, left, right, 1)
    ref = oracle_mod.advance(oracle_mod.Functional(kind=oracle_mod.OR_LHY), psi0, geo, stepper="RK4", dt=5e-4,
                             imaginary=False, nsteps=100)
    ctx = _ctx(C, shape, left, right)
    psi = ctx.field(psi0)
    f = C.functional_desc(C.GP5_FUNC_LHY)
    s = C.stepper_desc(C.GP5_STEP_RK4, dt=5e-4, imaginary=False)
    norms, ens = C.advance(ctx, f, s, psi, 100, log=True)
    assert rel_l2(psi.download_complex(), ref) <= 1e-12
    # conserved norm and energy in real time
    assert np.max(np.abs(norms[:, 0] - 3000.0)) / 3000.0 < 1e-9
    assert np.max(np.abs(ens - ens[0])) / abs(ens[0]) < 1e-8


def test_errors_surface(C):
    ctx = _ctx(C, (8, 8, 8), (-1,) * 3, (1,) * 3)
    other = _ctx(C, (8, 8, 4), (-1,) * 3, (1,) * 3)
    a, b = ctx.field(), other.field()
    with pytest.raises(C.GP5Error) as e:
        C.laplacian(ctx, a, b)
    assert e.value.code == C.GP5_ERR_ARG
    with pytest.raises(C.GP5Error):
        C.fft(ctx, a, a, 7)
    bad = C.stepper_desc(C.GP5_STEP_RK4, dt=1e-3, imaginary=True)
    bad.kind = 9
    with pytest.raises(C.GP5Error):
        C.advance(ctx, C.functional_desc(C.GP5_FUNC_LHY), bad, a, 1)
    nan = np.full((8, 8, 8, 1), np.nan, dtype=np.complex128)
    f = ctx.field(nan)
    with pytest.raises(C.GP5Error) as e:
        C.advance(ctx, C.functional_desc(C.GP5_FUNC_LHY), C.stepper_desc(dt=1e-3, imaginary=True), f, 1)
    assert e.value.code == C.GP5_ERR_NONFINITE
