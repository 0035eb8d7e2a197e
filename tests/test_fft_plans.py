"""Register-plan variants of the FFT passes and the autotuner that picks them.

Every variant (stage order, values per thread, columns per block) of the 64-, 96-,
128-, 192- and 512-point passes computes the same unnormalised DFT
(src/fourierTransform.cpp:184-214) and, inside a step, the same RK4 update
(src/stepper.cpp:46-70): each is held to the oracle on the role it serves (x
lines with the fused epilogue, y columns, the fused z multiplier pass).  The
autotuner (gp5_batch_tune) times the variants per batch class, reports its
choices, caches them per (architecture, grid) and reuses the cache."""
import numpy as np
import pytest

from conftest import rel_l2, seeded_field

capi = pytest.importorskip("gp5_amd.capi")

ROLE_AXIS = {0: "x", 1: "y", 2: "z"}
CASES = [(n, role, v) for n in (64, 96, 128, 192, 512) for role in (0, 1, 2) for v in (1, 2)]


@pytest.mark.gpu
@pytest.mark.parametrize("n,role,v", CASES, ids=[f"n{n}-{ROLE_AXIS[r]}-v{v}" for n, r, v in CASES])
def test_plan_variant_matches_oracle(oracle_mod, n, role, v):
    O = oracle_mod
    shape = [8, 12, 10]
    shape[role] = n
    shape = tuple(shape)
    left, right = (-4.0,) * 3, (4.0,) * 3
    print(f"plan variant n={n} role={role} v={v} shape={shape}", flush=True)
    ctx = capi.Context(shape, left, right, 1)
    vars_, counts = ctx.fft_plans()
    assert counts[role] == 3 and vars_ == [0, 0, 0]
    want = [0, 0, 0]
    want[role] = v
    ctx.set_fft_plans(want)
    assert ctx.fft_plans()[0] == want
    a = seeded_field(shape, 1, 71 + n + role)
    src, dst = ctx.field(a), ctx.field()
    for direction, sign in ((capi.GP5_FORWARD, -1), (capi.GP5_BACKWARD, +1)):
        capi.fft(ctx, src, dst, direction)
        assert rel_l2(dst.download_complex(), O.fft3d(a, sign)) <= 1e-13
    from gp5_amd import synth

    g = np.exp(-synth.r2_grid(shape, left, right) / 3.0)[..., None] * (1 + 0.05 * a)
    g = np.asfortranarray(g)
    geo = O.Geometry(shape, left, right, 1)
    # explicit RK4 needs dt k_max^2 / 2 inside its stability range: 0.6 here
    kmax = np.pi * n / 8.0
    dt = min(2e-4, 1.2 / kmax ** 2)
    for imag, renorm in ((False, None), (True, [2.0])):
        ref = O.advance(O.Functional(kind=O.OR_LHY), g, geo, stepper="RK4", dt=dt, imaginary=imag, nsteps=6,
                        renormalize=renorm)
        psi = ctx.field(g)
        s = capi.stepper_desc(capi.GP5_STEP_RK4, dt=dt, imaginary=imag, renormalize=renorm)
        capi.advance(ctx, capi.functional_desc(capi.GP5_FUNC_LHY), s, psi, 6)
        assert rel_l2(psi.download_complex(), ref) <= 1e-12


@pytest.mark.gpu
def test_plan_variants_rejected_out_of_range():
    ctx = capi.Context((64, 12, 10), (-4.0,) * 3, (4.0,) * 3, 1)
    with pytest.raises(capi.GP5Error):
        ctx.set_fft_plans([3, 0, 0])
    with pytest.raises(capi.GP5Error):
        ctx.set_fft_plans([0, 1, 0])  # the 12-point y axis runs on the generic kernels


def _gauss(n, L, sigma):
    x = -0.5 * L + (np.arange(n) + 0.5) * (L / n)
    r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
    return np.exp(-r2 / (2 * sigma * sigma)).astype(np.complex128)


@pytest.mark.gpu
def test_tuner_picks_register_plans_and_caches_them(oracle_mod, tmp_path, monkeypatch):
    O = oracle_mod
    cache = tmp_path / "plans.json"
    monkeypatch.setenv("GP5_TUNE_CACHE", str(cache))
    n = [64, 96, 64, 96, 64, 96]
    L = [22.4, 33.6, 22.4, 33.6, 22.4, 33.6]
    N = [60.0, 300.0, 90.0, 250.0, 120.0, 200.0]
    psi0 = [_gauss(n[i], L[i], 2.0 + 0.2 * i) for i in range(len(n))]
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    b = capi.Batch(n, L, N)
    for i, a in enumerate(psi0):
        b.upload(i, a)
    rep = b.tune(f, 0.1, 2)
    print(rep, flush=True)
    assert rep.count("class n=") == 2
    assert rep.count("x plan") == 2 and rep.count("z plan") == 2 and "tried" in rep
    assert cache.exists()
    text = cache.read_text()
    assert "/64\"" in text and "/96\"" in text
    b.advance(f, 0.1, 8)
    for i in range(len(n)):
        dx = L[i] / n[i]
        geo = O.Geometry((n[i],) * 3, (-L[i] / 2,) * 3, (L[i] / 2,) * 3, 1)
        ref = O.advance(O.Functional(kind=O.OR_LHY), psi0[i][..., None], geo, stepper="RK4", dt=0.1 * dx * dx,
                        imaginary=True, nsteps=8, renormalize=[N[i]])
        assert rel_l2(b.download(i)[..., None], ref) < 1e-12, i
    b.close()
    # a second batch of the same grids takes the cached plans without timing them
    b2 = capi.Batch(n, L, N)
    for i, a in enumerate(psi0):
        b2.upload(i, a)
    rep2 = b2.tune(f, 0.1, 2)
    print(rep2, flush=True)
    assert rep2.count("cached") == 2 and "tried" not in rep2
    b2.advance(f, 0.1, 8)
    for i in range(len(n)):
        dx = L[i] / n[i]
        geo = O.Geometry((n[i],) * 3, (-L[i] / 2,) * 3, (L[i] / 2,) * 3, 1)
        ref = O.advance(O.Functional(kind=O.OR_LHY), psi0[i][..., None], geo, stepper="RK4", dt=0.1 * dx * dx,
                        imaginary=True, nsteps=8, renormalize=[N[i]])
        assert rel_l2(b2.download(i)[..., None], ref) < 1e-12, i
    b2.close()
