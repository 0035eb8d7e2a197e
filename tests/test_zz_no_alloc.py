"""gp5_advance_noalloc (the entry point that allocates nothing) against gp5_advance.

Both run the same kernels on the same inputs, so the fields must be identical.
Two calls in a row, the second on a smaller input (another context), then the
rejection of an input beyond the static scratch."""
import numpy as np
import pytest

from conftest import seeded_field

pytestmark = pytest.mark.gpu


def _pair(C, shape, left, right, nc, a):
    ctx = C.Context(shape, left, right, nc)
    return ctx, ctx.field(a), ctx.field(a)


def test_noalloc_equals_advance_twice_then_smaller():
    from gp5_amd import capi as C
    from gp5_amd import synth

    # first call: 64^3 LHY droplet, real-time RK4, 7 steps (odd: the state ends in the scratch
    # output field and is copied home)
    shape, left, right, psi0 = synth.droplet_256(64)
    print("noalloc 64^3 RK4 x7", flush=True)
    ctx, p1, p2 = _pair(C, shape, left, right, 1, psi0)
    f = C.functional_desc(C.GP5_FUNC_LHY)
    s = C.stepper_desc(C.GP5_STEP_RK4, dt=5e-4, imaginary=False)
    C.advance(ctx, f, s, p1, 7)
    C.advance_noalloc(ctx, f, s, p2, 7)
    np.testing.assert_array_equal(p2.download_complex(), p1.download_complex())
    # the same context keeps working on its own buffers afterwards
    C.advance(ctx, f, s, p1, 2)
    C.advance_noalloc(ctx, f, s, p2, 2)
    np.testing.assert_array_equal(p2.download_complex(), p1.download_complex())

    # second call, smaller input: two coupled components on 24 x 20 x 16, imaginary-time
    # RK4 with per-stage renormalisation and a phase constraint, 5 steps
    shape2, left2, right2 = (24, 20, 16), (-5.0, -5.0, -4.0), (5.0, 5.0, 4.0)
    a = seeded_field(shape2, 2, 61, 0.5)
    x, y, _ = synth.axes(shape2, left2, right2)
    phase = np.empty(shape2 + (2,), order="F")
    phase[..., 0] = np.arctan2(y[None, :, None], x[:, None, None] + 0.01)
    phase[..., 1] = -phase[..., 0]
    print("noalloc 24x20x16x2 RK4 renorm phase x5", flush=True)
    ctx2, q1, q2 = _pair(C, shape2, left2, right2, 2, a)
    pb = ctx2.real(phase)
    f2 = C.functional_desc(C.GP5_FUNC_GP, couplings=[[1.0, 0.4], [0.4, 0.7]], masses=(1.0, 2.0))
    s2 = C.stepper_desc(C.GP5_STEP_RK4, dt=2e-3, imaginary=True, renormalize=[1.5, 0.5], phase=pb)
    C.advance(ctx2, f2, s2, q1, 5)
    C.advance_noalloc(ctx2, f2, s2, q2, 5)
    np.testing.assert_array_equal(q2.download_complex(), q1.download_complex())
    # and Euler (fewer scratch fields), real time
    s3 = C.stepper_desc(C.GP5_STEP_EULERO, dt=1e-3, imaginary=False)
    C.advance(ctx2, f2, s3, q1, 3)
    C.advance_noalloc(ctx2, f2, s3, q2, 3)
    np.testing.assert_array_equal(q2.download_complex(), q1.download_complex())


def test_noalloc_rejects_inputs_beyond_the_static_scratch():
    from gp5_amd import capi as C

    # RK4 on 128^3 x 2 components needs 5 x 64 MiB of fields: more than the 200 MiB
    shape, left, right = (128, 128, 128), (-8.0,) * 3, (8.0,) * 3
    ctx = C.Context(shape, left, right, 2)
    psi = ctx.field()
    f = C.functional_desc(C.GP5_FUNC_LHY)
    s = C.stepper_desc(C.GP5_STEP_RK4, dt=1e-4, imaginary=False)
    with pytest.raises(C.GP5Error) as e:
        C.advance_noalloc(ctx, f, s, psi, 1)
    assert e.value.code == C.GP5_ERR_UNSUPPORTED
    assert "static scratch holds" in str(e.value)
