"""The learned warm start (gp5_amd/surrogate.py with the committed trained weights):
what the network gains over the method without it.  Its seed of an untrapped
droplet is closer to the converged ground state than the cold Gaussian seed and the
flat-top baseline it corrects, and imaginary time from it is ahead of the cold start
at every checkpoint (energy error) up to convergence."""
import os

import numpy as np
import pytest

S = pytest.importorskip("gp5_amd.surrogate")


def test_scaled_problem_constants():
    # zero pressure of -3 rho + b rho^{3/2} at rho0 = (2.5 / b)^2; norm of the scaled field 4 pi / 3
    for b in (2.0, 2.5, 3.0):
        r0 = S.rho0(b)
        assert 0.5 * -3.0 * r0 ** 2 + 0.6 * b * r0 ** 2.5 == pytest.approx(0.0, abs=1e-12)
    N, b = 700.0, 2.2
    R0 = S.radius(N, b)
    assert S.rho0(b) * 4.0 / 3.0 * np.pi * R0 ** 3 == pytest.approx(N, rel=1e-12)


def test_symmetrize_and_radial_profile():
    rng = np.random.default_rng(0)
    phi = rng.standard_normal((S.GRID,) * 3)
    s = S.symmetrize(phi, (0.1, 0.1, 0.05))
    assert np.allclose(s, s[::-1]) and np.allclose(s, s[:, :, ::-1])
    assert np.allclose(s, np.transpose(s, (1, 0, 2)))
    ru, prof = S.radial_profile(S.baseline(500.0, 2.5))
    assert prof[0] == pytest.approx(1.0, abs=1e-4) and prof[-1] < 1e-6
    assert np.all(np.diff(ru) > 0)


@pytest.mark.gpu
def test_seed_beats_cold_and_baseline(oracle_mod):
    from gp5_amd import capi, synth

    assert os.path.exists(S.DEFAULT_WEIGHTS)
    sur = S.Surrogate()
    n, L, N, psi0 = synth.droplet_batch(64, 3)
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    for i in (5, 18):  # 64^3 and 96^3 instances
        dx = L[i] / n[i]
        ctx = capi.Context((n[i],) * 3, (-L[i] / 2,) * 3, (L[i] / 2,) * 3, 1)
        s = capi.stepper_desc(capi.GP5_STEP_RK4, dt=0.1 * dx * dx, imaginary=True, renormalize=[N[i]])
        seeds = {"cold": psi0[i], "net": sur.seed(N[i], (0, 0, 0), 2.5, n[i], L[i]),
                 "base": S.sqrt_rho_interp(S.baseline(N[i], 2.5), N[i], 2.5, n[i], L[i])}
        E, fields = {}, {}
        for k, a in seeds.items():
            psi = ctx.field(a)
            _, E[k] = capi.advance(ctx, f, s, psi, 1200, log=True)
            fields[k] = psi.download_complex()
        ref = fields["cold"]  # converged to ~1e-10 by 1200 steps
        Ef = E["cold"][-1]
        err = {k: np.abs(e - Ef) / abs(Ef) for k, e in E.items()}
        # the network's seed: 10x closer in energy than the cold seed, 5x than its baseline
        assert err["net"][0] < 0.1 * err["cold"][0]
        assert err["net"][0] < 0.2 * err["base"][0]
        # and imaginary time from it stays ahead through the transient
        for t in (100, 300):
            assert err["net"][t - 1] < err["cold"][t - 1], (i, t)
        assert np.linalg.norm(np.abs(fields["net"]) - np.abs(ref)) < 1e-3 * np.linalg.norm(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("N,w,b", [(300.0, (0.0, 0.0, 0.0), 2.5), (2345.0, (0.1, 0.05, 0.17), 2.2)])
def test_device_features_match_host_features(N, w, b):
    """gp5_nn_droplet_features reproduces features() (float32, rounded to bf16): the
    trap, radius and constant channels bit for bit, the tanh baseline within one bf16
    step; the device-built prediction matches the host-built one."""
    from gp5_amd import nn

    if not os.path.exists(S.DEFAULT_WEIGHTS):
        pytest.skip("no trained weights")
    sur = S.Surrogate()
    host = S.features(N, w, b)
    want = nn.to_bf16(host).reshape(-1, 16)
    R0 = S.radius(N, b)
    a = [(wi / 0.2) ** 2 / 18.0 for wi in w]
    consts = [np.log(N / 50.0) / np.log(100.0), 2.0 * (b - 2.5), w[0] / 0.2, w[1] / 0.2, w[2] / 0.2]
    sur.net.set_droplet_features(np.float32(R0 / S.SURFACE), np.float32(a), np.float32(consts), S.U_BOX)
    got = sur.net.t["X"].download(np.uint16, (S.GRID ** 3 * 16,)).reshape(-1, 16)
    assert np.array_equal(got[:, 1:], want[:, 1:])
    # the baseline: tanhf and numpy's float32 tanh may differ in the last place, which
    # shows as one bf16 step, or as 3e-8 against 0 far outside the droplet
    # (1 - tanh cancels: a few float32 ulps of tanh near 1 are ~2e-7 absolute)
    g0, w0 = nn.from_bf16(got[:, 0]), nn.from_bf16(want[:, 0])
    excess = np.abs(g0 - w0) - 2.0 ** -7 * np.abs(w0)
    print(f"baseline channel: max |diff| {np.abs(g0 - w0).max():.3e}, max excess {excess.max():.3e}", flush=True)
    assert excess.max() <= 5e-7
    # where 1 - tanh does not cancel, a last-place tanh difference stays below half a bf16 step
    core = np.abs(w0) > 1e-2
    assert (got[core, 0] != want[core, 0]).mean() < 1e-2
    base = sur.net.t["base"].download(np.float32, (S.GRID ** 3,))
    assert np.allclose(base, host[..., 0].ravel(), rtol=0, atol=2e-7)
    dev = sur.predict(N, w, b, device_features=True)
    hst = sur.predict(N, w, b, device_features=False)
    assert np.linalg.norm(dev - hst) <= 1e-3 * np.linalg.norm(hst)


@pytest.mark.gpu
def test_batched_predictions_match_single():
    """predict_many / seed_many with a batch-4 network (6 droplets: one full and one
    partial forward pass) give the batch-1 predictions to the bf16 level (the kernels
    may sum in another order at another batch size), and the seeds agree likewise."""
    sur = S.Surrogate(batch=4)
    items = [(60.0, (0.0, 0.0, 0.0), 2.5), (300.0, (0.0, 0.0, 0.0), 2.5), (1500.0, (0.0, 0.0, 0.0), 2.5),
             (4000.0, (0.0, 0.0, 0.0), 2.5), (800.0, (0.1, 0.1, 0.05), 2.2), (120.0, (0.0, 0.0, 0.0), 2.8)]
    many = sur.predict_many(items)
    for (N, w, b), got in zip(items, many):
        one = sur.predict(N, w, b)
        assert got.shape == one.shape
        assert np.max(np.abs(got - one)) <= 2e-2 * np.max(np.abs(one)), (N, w, b)
    seeds = sur.seed_many([(N, w, b, 32, 24.0) for N, w, b in items])
    for (N, w, b), psi in zip(items, seeds):
        ref = sur.seed(N, w, b, 32, 24.0)
        assert np.max(np.abs(psi - ref)) <= 2e-2 * np.max(np.abs(ref)), (N, w, b)


@pytest.mark.gpu
def test_network_coarse_starts_beat_gaussian_coarse_starts():
    """What the trained network gains over the method without it: two-level ground
    states (coarse grid n/2, then n) of the same droplets, coarse level started from the
    network's seeds or from Gaussians (gp5_amd.ground_state.two_level).  The network's
    coarse level takes fewer imaginary-time steps in total, and both reach the same
    fine-grid energies within the convergence tolerance."""
    from gp5_amd import capi, ground_state, synth

    assert os.path.exists(S.DEFAULT_WEIGHTS)
    n, L, N, _ = synth.droplet_batch(64, 3)
    pick = [i for i in range(len(n)) if n[i] in (64, 96)][:12]
    n, L, N = [n[i] for i in pick], [L[i] for i in pick], [N[i] for i in pick]
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    sur = S.Surrogate(batch=4)
    runs = {}
    for name, kw in (("gaussian", {}), ("network", {"seed_many": lambda items: sur.seed_many(
            [(Nk, (0.0, 0.0, 0.0), 2.5, k, Lk) for _, k, Lk, Nk in items])})):
        fb, rep = ground_state.two_level(n, L, N, f, tol=1e-10, tune=False, **kw)
        fb.close()
        runs[name] = rep
        print(name, "coarse steps", int(np.abs(rep["coarse_steps"]).sum()), "fine steps",
              int(np.abs(rep["steps"]).sum()), f"{rep['total_s']:.2f} s", flush=True)
    g, w = runs["gaussian"], runs["network"]
    assert np.all(g["steps"] > 0) and np.all(w["steps"] > 0)
    assert np.abs(w["coarse_steps"]).sum() < np.abs(g["coarse_steps"]).sum()
    assert np.all(np.abs(w["energy"] - g["energy"]) <= 5e-10 * np.abs(g["energy"]))
