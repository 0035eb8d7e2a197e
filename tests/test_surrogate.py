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
