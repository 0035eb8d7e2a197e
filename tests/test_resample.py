"""Spectral resampling between grids of one box (gp5_resample, gp5_batch_resample).

The CPU reference (oracle/resample_ref.py) evaluates the trigonometric
interpolant directly; it is pinned by band-limited functions, which every
resampling to a grid that resolves them must reproduce exactly, and by the
identity at equal sizes.  The GPU path must match the reference."""
import numpy as np
import pytest

from conftest import rel_l2

from oracle import resample_ref as R


def _waves(shape, L, modes, seed=0):
    """sum of plane waves exp(2 pi i k.x / L) with random complex amplitudes, modes < Nyquist."""
    rng = np.random.default_rng(seed)
    x = [R.centres(n, L) for n in shape]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    out = np.zeros(shape, complex)
    for k in modes:
        amp = rng.standard_normal() + 1j * rng.standard_normal()
        out += amp * np.exp(2j * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) / L)
    return out


MODES = [(0, 0, 0), (1, -2, 3), (-3, 4, -1), (4, 0, -4), (-2, -3, 2)]


@pytest.mark.parametrize("src,dst", [((10, 12, 14), (20, 24, 28)), ((11, 10, 9), (16, 15, 13)),
                                     ((20, 18, 16), (10, 11, 12)), ((12, 12, 12), (12, 12, 12))])
def test_reference_reproduces_band_limited_functions(src, dst):
    L = 7.5
    a = _waves(src, L, MODES)
    got = R.resample(a, L, dst)[..., 0]
    assert np.abs(got - _waves(dst, L, MODES)).max() < 1e-12


def test_reference_identity_and_round_trip():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((8, 7, 6)) + 1j * rng.standard_normal((8, 7, 6))
    assert np.abs(R.resample(a, 3.0, (8, 7, 6))[..., 0] - a).max() < 1e-12
    up = R.resample(a, 3.0, (16, 15, 12))
    back = R.resample(up, 3.0, (8, 7, 6))
    assert np.abs(back[..., 0] - a).max() < 1e-12


def test_reference_nyquist_cosine_is_even():
    """The even-size Nyquist mode is taken as a cosine: a real sample set stays real."""
    rng = np.random.default_rng(2)
    a = rng.standard_normal((8, 8, 8))
    out = R.resample(a, 2.0, (12, 10, 16))
    assert np.abs(out.imag).max() < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("src,dst,nc", [((16, 12, 10), (32, 24, 20), 1), ((15, 10, 9), (30, 16, 12), 2),
                                        ((32, 32, 32), (16, 24, 20), 1), ((24, 24, 24), (24, 24, 24), 1),
                                        ((48, 48, 48), (96, 96, 96), 1)])
def test_gpu_resample_matches_reference(src, dst, nc):
    from gp5_amd import capi

    L = 9.0
    rng = np.random.default_rng(sum(src) + sum(dst))
    a = np.asfortranarray(rng.standard_normal(src + (nc,)) + 1j * rng.standard_normal(src + (nc,)))
    print(f"resample {src} -> {dst} nc={nc}", flush=True)
    cs = capi.Context(src, [-L / 2] * 3, [L / 2] * 3, nc)
    cd = capi.Context(dst, [-L / 2] * 3, [L / 2] * 3, nc)
    bs, bd = cs.field(a), cd.field()
    capi.resample(cs, bs, cd, bd)
    got = bd.download_complex()
    want = R.resample(a, L, dst)
    assert rel_l2(got, want) < 1e-13
    w = _waves(src, L, MODES, seed=3)
    bs.upload_complex(np.repeat(w[..., None], nc, axis=3))
    capi.resample(cs, bs, cd, bd)
    if all(d >= 9 for d in dst):
        assert np.abs(bd.download_complex()[..., 0] - _waves(dst, L, MODES, seed=3)).max() < 1e-11


@pytest.mark.gpu
def test_gpu_resample_rejects_other_boxes_and_component_counts():
    from gp5_amd import capi

    cs = capi.Context((8, 8, 8), [-1] * 3, [1] * 3, 1)
    cd = capi.Context((16, 16, 16), [-1] * 3, [1.5] * 3, 1)
    with pytest.raises(capi.GP5Error, match="same box"):
        capi.resample(cs, cs.field(), cd, cd.field())
    c2 = capi.Context((16, 16, 16), [-1] * 3, [1] * 3, 2)
    with pytest.raises(capi.GP5Error, match="components"):
        capi.resample(cs, cs.field(), c2, c2.field())


@pytest.mark.gpu
def test_gpu_batch_resample_between_batches():
    from gp5_amd import capi

    L = [11.2, 16.8, 11.2]
    coarse = capi.Batch([16, 24, 32], L, [10.0, 20.0, 30.0])
    fine = capi.Batch([32, 48, 48], L, [10.0, 20.0, 30.0])
    fields = []
    for i, n in enumerate(coarse.n):
        w = _waves((n,) * 3, L[i], MODES, seed=10 + i)
        fields.append(w)
        coarse.upload(i, w)
    for i in range(3):
        fine.resample_from(i, coarse, i)
    for i, n in enumerate(fine.n):
        assert np.abs(fine.download(i) - _waves((n,) * 3, L[i], MODES, seed=10 + i)).max() < 1e-11
    bad = capi.Batch([32], [12.0], [10.0])
    with pytest.raises(capi.GP5Error, match="boxes differ"):
        bad.resample_from(0, coarse, 0)
    coarse.close(), fine.close(), bad.close()


@pytest.mark.gpu
def test_two_level_ground_states_match_cold_starts():
    """Ground states started on the half-resolution grid and finished on the full one
    reach the cold-start energies (to the convergence tolerance) in fewer fine steps."""
    from gp5_amd import capi, ground_state

    N = [120.0, 400.0, 900.0]
    n = [64, 96, 96]
    L = [k * 0.35 for k in n]
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    cold = capi.Batch(n, L, N)
    for i in range(3):
        cold.upload(i, ground_state.gaussian_seed(n[i], L[i], N[i]))
    steps_c, en_c, _ = cold.ground_state(f, 0.1, 1e-10, 100, 100000)
    cold.close()
    fb, rep = ground_state.two_level(n, L, N, f, tol=1e-10, tune=False)
    assert rep["coarse_grids"] == [32, 48, 48]
    assert (rep["steps"] > 0).all() and (steps_c > 0).all()
    assert np.max(np.abs(rep["energy"] - en_c) / np.abs(en_c)) < 1e-9
    assert rep["steps"].sum() * 3 <= steps_c.sum(), (rep["steps"], steps_c)
    psi = fb.download(1)
    dV = (L[1] / n[1]) ** 3
    assert abs((np.abs(psi) ** 2).sum() * dV - N[1]) < 1e-8 * N[1]
    fb.close()
