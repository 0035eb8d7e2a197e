"""Many-component contexts and batches of independent instances.

A context with more than two components holds uncoupled instances of one grid; a
batch groups instances of several grids (classes) and runs the reference's
imaginary-time RK4 with per-stage renormalisation (src/stepper.cpp:46-70) on all
of them, to ground states through hipGraph replays.  Each instance must match the
oracle run of that instance alone (the instances never interact)."""
import numpy as np
import pytest

from conftest import rel_l2, seeded_field

capi = pytest.importorskip("gp5_amd.capi")


def test_batch_binding_checks_lengths_on_host():
    with pytest.raises(capi.GP5Error):
        capi.Batch([16, 16], [10.0], [5.0, 5.0])


def _gauss(n, L, sigma, seed):
    x = -0.5 * L + (np.arange(n) + 0.5) * (L / n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    g = np.exp(-(X**2 + Y**2 + Z**2) / (2 * sigma**2))
    u = seeded_field((n, n, n), 1, seed, 0.02)[..., 0]
    return (g * (1 + u)).astype(np.complex128)


def _geo(O, n, L, nc=1):
    return O.Geometry((n, n, n), (-L / 2,) * 3, (L / 2,) * 3, nc)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["lhy", "gp"])
def test_many_component_context_matches_single_instances(oracle_mod, kind):
    """nc = 5 uncoupled components in one context: RK4 real time, 6 steps."""
    O = oracle_mod
    shape, left, right = (16, 12, 20), (-4.0, -3.0, -5.0), (4.0, 3.0, 5.0)
    nc = 5
    a0 = seeded_field(shape, nc, 21, 0.6)
    ctx = capi.Context(shape, left, right, nc)
    psi = ctx.field(a0)
    if kind == "lhy":
        f = capi.functional_desc(capi.GP5_FUNC_LHY)
        F = O.Functional(kind=O.OR_LHY)
    else:
        f = capi.functional_desc(capi.GP5_FUNC_GP, couplings=np.array([[0.7]]))
        F = O.Functional(kind=O.OR_GP, couplings=np.array([[0.7]]))
    s = capi.stepper_desc(capi.GP5_STEP_RK4, dt=2e-3, imaginary=False)
    norms, ens = capi.advance(ctx, f, s, psi, 6, log=True)
    out = psi.download_complex()
    geo1 = O.Geometry(shape, left, right, 1)
    e0 = 0.0
    for c in range(nc):
        ref = O.advance(F, a0[..., c:c + 1], geo1, stepper="RK4", dt=2e-3, imaginary=False, nsteps=6)
        assert rel_l2(out[..., c:c + 1], ref) < 1e-12, c
        assert norms[0, c] == pytest.approx(O.norm(a0[..., c:c + 1], geo1, 0), rel=1e-12)
        e0 += O.energy(F, a0[..., c:c + 1], geo1)
    assert ens[0] == pytest.approx(e0, rel=1e-11)


@pytest.mark.gpu
def test_many_component_apply_and_energy(oracle_mod):
    O = oracle_mod
    shape, left, right = (24, 20, 16), (-5.0, -5.0, -4.0), (5.0, 5.0, 4.0)
    nc = 7
    a0 = seeded_field(shape, nc, 22, 0.5)
    ctx = capi.Context(shape, left, right, nc)
    src, dst = ctx.field(a0), ctx.field()
    f = capi.functional_desc(capi.GP5_FUNC_LHY, lhy_a=-2.5, lhy_b=2.0)
    capi.functional_apply(ctx, f, src, dst)
    out = dst.download_complex()
    F = O.Functional(kind=O.OR_LHY, lhy_a=-2.5, lhy_b=2.0)
    geo1 = O.Geometry(shape, left, right, 1)
    for c in range(nc):
        assert rel_l2(out[..., c:c + 1], O.apply(F, a0[..., c:c + 1], geo1)) < 1e-12
    e = capi.energy(ctx, f, src)
    assert e == pytest.approx(sum(O.energy(F, a0[..., c:c + 1], geo1) for c in range(nc)), rel=1e-11)
    for c in (0, 3, 6):
        assert capi.norm(ctx, src, c) == pytest.approx(O.norm(a0[..., c:c + 1], geo1, 0), rel=1e-12)


def _batch_case():
    # three classes: two 16^3 boxes of different size and one 24^3 (generic kernels)
    n = [16, 24, 16, 24, 16, 24, 16]
    L = [10.0, 14.0, 10.0, 14.0, 12.0, 14.0, 10.0]
    N = [60.0, 150.0, 90.0, 80.0, 120.0, 200.0, 50.0]
    sig = [1.6, 2.2, 1.8, 2.0, 2.0, 2.4, 1.5]
    psi0 = [_gauss(n[i], L[i], sig[i], 30 + i) for i in range(len(n))]
    return n, L, N, psi0


@pytest.mark.gpu
def test_batch_advance_matches_oracle_per_instance(oracle_mod):
    O = oracle_mod
    n, L, N, psi0 = _batch_case()
    b = capi.Batch(n, L, N)
    assert b.classes == 3
    assert b.class_of(0) == b.class_of(2) != b.class_of(4)
    for i, a in enumerate(psi0):
        b.upload(i, a)
    b.set_lhy(3, -2.6, 2.2)
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    dt_factor = 0.3
    b.advance(f, dt_factor, 10)
    for i in range(len(n)):
        dx = L[i] / n[i]
        F = O.Functional(kind=O.OR_LHY, lhy_a=-2.6 if i == 3 else -3.0, lhy_b=2.2 if i == 3 else 2.5)
        ref = O.advance(F, psi0[i][..., None], _geo(O, n[i], L[i]), stepper="RK4", dt=dt_factor * dx * dx,
                        imaginary=True, nsteps=10, renormalize=[N[i]])
        assert rel_l2(b.download(i)[..., None], ref) < 1e-12, i
    b.close()


@pytest.mark.gpu
def test_batch_potentials_and_tuned_plans_match_oracle(oracle_mod):
    """Per-instance traps, and the autotuner's plans and tiles keep results identical."""
    O = oracle_mod
    n, L, N, psi0 = _batch_case()
    b = capi.Batch(n, L, N)
    Vs = {}
    for i, a in enumerate(psi0):
        b.upload(i, a)
        if i % 2 == 0:
            x = -0.5 * L[i] + (np.arange(n[i]) + 0.5) * (L[i] / n[i])
            X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
            w = (0.1 + 0.02 * i, 0.15, 0.05)
            Vs[i] = 0.5 * (w[0] ** 2 * X**2 + w[1] ** 2 * Y**2 + w[2] ** 2 * Z**2)
            b.set_potential(i, Vs[i])
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    report = b.tune(f, 0.3, steps=2)
    assert report.count("class n=") == 3
    b.advance(f, 0.3, 6)
    for i in range(len(n)):
        dx = L[i] / n[i]
        V = Vs[i][..., None] if i in Vs else None
        ref = O.advance(O.Functional(kind=O.OR_LHY, V=V), psi0[i][..., None], _geo(O, n[i], L[i]), stepper="RK4",
                        dt=0.3 * dx * dx, imaginary=True, nsteps=6, renormalize=[N[i]])
        assert rel_l2(b.download(i)[..., None], ref) < 1e-12, i
    b.close()


@pytest.mark.gpu
def test_batch_graph_replays_match_eager_steps():
    """ground_state stopped by max_steps (graph replays of 8 steps) equals eager advance."""
    n, L, N, psi0 = _batch_case()
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    b1, b2 = capi.Batch(n, L, N), capi.Batch(n, L, N)
    for i, a in enumerate(psi0):
        b1.upload(i, a)
        b2.upload(i, a)
    steps, en, ms = b1.ground_state(f, 0.3, tol=1e-300, check_every=8, max_steps=24)
    assert np.all(steps == -24)
    b2.advance(f, 0.3, 8)
    b2.advance(f, 0.3, 8)
    b2.advance(f, 0.3, 8)
    for i in range(len(n)):
        assert rel_l2(b1.download(i), b2.download(i)) < 1e-14, i
    b1.close()
    b2.close()


@pytest.mark.gpu
def test_batch_ground_state_converges_and_matches_oracle(oracle_mod):
    """Trapped contact-interaction instances (gapped spectra: a few hundred steps), three
    classes that shrink as instances retire, traps and normalisations per instance; each retired field, its step count and
    energy are checked against the oracle run of that instance alone."""
    O = oracle_mod
    n = [16, 24, 16, 24, 16, 16, 24]
    L = [10.0, 12.0, 10.0, 12.0, 11.0, 10.0, 12.0]
    N = [1.0, 5.0, 2.0, 0.5, 3.0, 8.0, 1.5]
    w = [1.0, 0.9, 1.3, 1.1, 1.0, 1.5, 0.8]
    b = capi.Batch(n, L, N)
    psi0, V = [], []
    for i in range(len(n)):
        x = -0.5 * L[i] + (np.arange(n[i]) + 0.5) * (L[i] / n[i])
        X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
        V.append(0.5 * w[i] ** 2 * (X**2 + 0.9 * Y**2 + 1.2 * Z**2))
        psi0.append(_gauss(n[i], L[i], 1.4, 40 + i))
        b.upload(i, psi0[i])
        b.set_potential(i, V[i])
    f = capi.functional_desc(capi.GP5_FUNC_GP, couplings=np.array([[1.0]]))
    F_of = lambda i: O.Functional(kind=O.OR_GP, couplings=np.array([[1.0]]), V=V[i][..., None])  # noqa: E731
    # explicit RK4 in imaginary time is stable for dt * max(V + k^2/2) < 2.78: the
    # traps reach V ~ 40 at the box corners, so dt = 0.1 dx^2 (0.3 dx^2 is not)
    K, tol, dtf = 10, 1e-10, 0.1
    steps, en, ms = b.ground_state(f, dtf, tol=tol, check_every=K, max_steps=3000)
    print("steps", steps.tolist(), "energies", en.tolist(), "ms", ms, flush=True)
    assert np.all(steps > 0)
    assert np.all(steps % K == 0)
    assert len(set(steps.tolist())) > 1  # instances retire at different times
    for i in range(len(n)):
        dx = L[i] / n[i]
        geo = _geo(O, n[i], L[i])
        dt = dtf * dx * dx
        F = F_of(i)
        k = int(steps[i])
        a = O.advance(F, psi0[i][..., None], geo, stepper="RK4", dt=dt, imaginary=True, nsteps=k - K - 1,
                      renormalize=[N[i]])
        e_prev = O.energy(F, a, geo)  # energy logged at the start of the previous check's last step
        a = O.advance(F, a, geo, stepper="RK4", dt=dt, imaginary=True, nsteps=K, renormalize=[N[i]])
        e_last = O.energy(F, a, geo)
        assert en[i] == pytest.approx(e_last, rel=1e-10), i
        assert abs(e_last - e_prev) <= tol * abs(e_last) * (1 + 1e-6), i
        a = O.advance(F, a, geo, stepper="RK4", dt=dt, imaginary=True, nsteps=1, renormalize=[N[i]])
        assert rel_l2(b.download(i)[..., None], a) < 1e-11, i
    b.close()


@pytest.mark.gpu
def test_batch_error_keeps_the_slot_map_true():
    """A non-finite instance stops a ground-state run with GP5_ERR_NONFINITE naming it.
    Instances of the other class may retire (and move slots on the device) in the replay
    that runs while the host reads the error; afterwards every instance still downloads
    its own field (each normalised to its own N)."""
    n = [16, 16, 16, 24, 24]
    L = [10.0, 10.0, 10.0, 12.0, 12.0]
    N = [1.0, 2.0, 3.0, 4.0, 5.0]
    b = capi.Batch(n, L, N)
    for i in range(4):
        b.upload(i, _gauss(n[i], L[i], 1.5 + 0.1 * i, 50 + i))
    b.upload(4, np.full((24, 24, 24), np.nan, dtype=np.complex128))
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    with pytest.raises(capi.GP5Error) as e:
        b.ground_state(f, 0.1, tol=5e-2, check_every=10, max_steps=2000)
    assert e.value.code == capi.GP5_ERR_NONFINITE
    assert "instance 4" in str(e.value)
    for i in range(4):
        a = b.download(i)
        dV = (L[i] / n[i]) ** 3
        assert np.sum(np.abs(a) ** 2) * dV == pytest.approx(N[i], rel=1e-10), i
    b.close()
