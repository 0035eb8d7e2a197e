"""The large workload: two LHY droplets colliding in a 512^3 box (L = 64, real-time
RK4, dt = 2.5e-4), on the HIP path, against the oracle after one step (relative L2
<= 1e-12), with norm and energy drift over a short run."""
import pytest

from conftest import rel_l2


@pytest.mark.gpu
def test_collision_512_one_step_matches_oracle(oracle_mod):
    from gp5_amd import capi, synth

    O = oracle_mod
    O.set_threads(16)
    shape, left, right, psi0, q = synth.droplet_collision(512)
    print(f"collision 512^3 q={q:.4f}", flush=True)
    ctx = capi.Context(shape, left, right, 1)
    psi = ctx.field(psi0)
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    s = capi.stepper_desc(capi.GP5_STEP_RK4, dt=2.5e-4, imaginary=False)
    capi.advance(ctx, f, s, psi, 1)
    got = psi.download_complex()
    psi.free()
    ctx.close()
    geo = O.Geometry(shape, left, right, 1)
    ref = O.advance(O.Functional(kind=O.OR_LHY), psi0, geo, stepper="RK4", dt=2.5e-4, imaginary=False, nsteps=1)
    err = rel_l2(got, ref)
    print(f"relative L2 after one step: {err:.2e}", flush=True)
    assert err <= 1e-12


@pytest.mark.gpu
def test_collision_512_conserves_norm_and_energy():
    from gp5_amd import capi, synth

    shape, left, right, psi0, q = synth.droplet_collision(512)
    ctx = capi.Context(shape, left, right, 1)
    psi = ctx.field(psi0)
    del psi0
    f = capi.functional_desc(capi.GP5_FUNC_LHY)
    s = capi.stepper_desc(capi.GP5_STEP_RK4, dt=2.5e-4, imaginary=False)
    norms, ens = capi.advance(ctx, f, s, psi, 40, log=True)
    psi.free()
    ctx.close()
    assert abs(norms[0, 0] - 8000.0) < 1e-9 * 8000.0
    assert abs(norms[-1, 0] - norms[0, 0]) < 1e-9 * norms[0, 0]
    assert abs(ens[-1] - ens[0]) < 1e-7 * abs(ens[0])
