"""Pins the CPU oracle (oracle/gp5_oracle.c) before it is trusted as the checker.

Every oracle function is checked against something that is not a retyping of
its own formula:
  * the reference's own known-answer tests (tst/fft_test.cpp:10-153,
    python/test/test.py:92-147) at their shapes and tolerances, and tighter;
  * brute-force O(M^2) DFTs on tiny grids and numpy's pocketfft on the
    reference's test shapes (FFTW's unnormalised c2c convention);
  * closed forms: plane waves are Laplacian eigenfunctions, Gaussian norms,
    the harmonic-oscillator ground-state energy, Euler/RK4 amplification
    polynomials on a plane wave;
  * the nonlinear branches (contact g|psi|^2 psi, the two-component couplings
    with masses, the unlocked LHY form, the LHY coefficients, uniform
    potentials) on plane waves, where every stage is a multiple of the wave and
    a 3-D run reduces to a scalar recurrence on the amplitudes, and energies of
    plane waves;
  * an independent numpy restatement of the steppers (tests/np_ref.py) for
    every functional, component count, renormalisation and constraint.
"""
import numpy as np
import pytest

import np_ref
from conftest import rel_l2, seeded_field


def _dft_matrix(n, sign):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)


@pytest.mark.parametrize("shape", [(4, 6, 5), (3, 7, 2), (8, 9, 12), (11, 1, 13)])
def test_fft_brute_force(oracle_mod, shape):
    a = seeded_field(shape, 1, seed=3)
    for sign in (-1, 1):
        # X[k1,k2,k3] = sum_n x[n] e^{sign 2 pi i k.n/N}, via explicit matrices
        ref = np.einsum("ai,bj,ck,ijk->abc", _dft_matrix(shape[0], sign), _dft_matrix(shape[1], sign),
                        _dft_matrix(shape[2], sign), a[..., 0])
        got = oracle_mod.fft3d(a, sign)[..., 0]
        assert rel_l2(got, ref) < 1e-14


@pytest.mark.parametrize("shape,nc", [((50, 170, 150), 1), ((30, 30, 30), 2), ((64, 96, 20), 1), ((7, 11, 17), 1)])
def test_fft_matches_pocketfft(oracle_mod, shape, nc):
    a = seeded_field(shape, nc, seed=7)
    fwd = oracle_mod.fft3d(a, -1)
    bwd = oracle_mod.fft3d(a, +1)
    assert rel_l2(fwd, np.fft.fftn(a, axes=(0, 1, 2))) < 1e-14
    assert rel_l2(bwd, np.fft.ifftn(a, axes=(0, 1, 2)) * np.prod(shape)) < 1e-14


def test_fft_forward_backward_reference(oracle_mod):
    """tst/fft_test.cpp:10-83 (50x170x150, Gaussian sigma = (1, 0.5, 0.2), TOL 1e-6)."""
    from gp5_amd import synth

    shape, left, right = (50, 170, 150), (-0.5,) * 3, (0.5,) * 3
    x, y, z = synth.axes(shape, left, right)
    g = np.exp(-(x[:, None, None] ** 2 / 2 + y[None, :, None] ** 2 / (2 * 0.25) + z[None, None, :] ** 2 / (2 * 0.04)))
    a = np.asfortranarray(g[..., None].astype(np.complex128))
    back = oracle_mod.fft3d(oracle_mod.fft3d(a, -1), +1) / np.prod(shape)
    err = np.max(np.abs(back - a))
    assert err <= 1e-6
    assert err <= 1e-14


def test_laplacian_closed_form_reference(oracle_mod):
    """tst/fft_test.cpp:85-153 (300x250x77, sigma = 0.05, max abs error <= 1e-6)."""
    from gp5_amd import synth

    shape, left, right = (300, 250, 77), (-0.5,) * 3, (0.5,) * 3
    alpha = 1 / (2 * 0.05 ** 2)
    r2 = synth.r2_grid(shape, left, right)
    a = np.asfortranarray(np.exp(-alpha * r2)[..., None].astype(np.complex128))
    L = oracle_mod.laplacian(a, oracle_mod.Geometry(shape, left, right, 1))[..., 0]
    expect = np.exp(-alpha * r2) * (3 - 2 * alpha * r2) * (-2 * alpha)
    err = np.max(np.abs(L - expect))
    assert err <= 1e-6
    assert err <= 1e-9 * np.max(np.abs(expect))


@pytest.mark.parametrize("shape", [(8, 6, 10), (9, 7, 5), (16, 1, 4)])
def test_plane_waves_are_eigenfunctions(oracle_mod, shape):
    """lap e^{i k.x} = -|k|^2 e^{i k.x} for every grid wavenumber, incl. Nyquist and odd N."""
    left, right = (-1.0, -2.0, -0.5), (2.0, 1.5, 3.0)
    geo = oracle_mod.Geometry(shape, left, right, 1)
    from gp5_amd import synth

    x, y, z = synth.axes(shape, left, right)
    L = [right[d] - left[d] for d in range(3)]
    for m in [(0, 0, 0), (1, 2, 1), (shape[0] // 2, 0, 0), (-1, shape[1] // 2, shape[2] // 2), (2, -1, -2)]:
        # representable wavenumbers only: m_d reduced modulo N_d (a 1-point axis carries k = 0)
        m = [((m[d] + shape[d] // 2) % shape[d]) - shape[d] // 2 for d in range(3)]
        k = [2 * np.pi * m[d] / L[d] for d in range(3)]
        f = np.exp(1j * (k[0] * x[:, None, None] + k[1] * y[None, :, None] + k[2] * z[None, None, :]))
        lap = oracle_mod.laplacian(np.asfortranarray(f[..., None]), geo)[..., 0]
        # the reference's k for the Nyquist mode is -N/2 dk; |k|^2 is the same
        expect = -(k[0] ** 2 + k[1] ** 2 + k[2] ** 2) * f
        assert np.max(np.abs(lap - expect)) < 1e-11 * max(1.0, np.max(np.abs(expect)))


@pytest.mark.parametrize("n", [1, 2, 5, 8, 9, 256])
def test_momentum_rule(oracle_mod, n):
    """geometry.cpp:127-148 equals numpy's fftfreq ordering times 2 pi / L."""
    dk = 0.37
    got = np.array([oracle_mod.momentum(i, n, dk) for i in range(n)])
    expect = np.fft.fftfreq(n, d=1.0 / n) * dk
    np.testing.assert_array_equal(got, expect)


def test_lhy_apply_closed_form_reference(oracle_mod):
    """python/test/test.py:92-119: H e^{-r^2} on 100^3 over [-4,4]^3 within 1e-4."""
    from gp5_amd import synth

    shape, left, right = (100, 100, 100), (-4.0,) * 3, (4.0,) * 3
    r2 = synth.r2_grid(shape, left, right)
    psi = np.exp(-r2)
    out = oracle_mod.apply(oracle_mod.Functional(kind=oracle_mod.OR_LHY),
                           np.asfortranarray(psi[..., None].astype(np.complex128)),
                           oracle_mod.Geometry(shape, left, right, 1))[..., 0]
    expect = (3 - 2 * r2) * psi + (-3 * psi ** 2 + 2.5 * psi ** 3) * psi
    assert np.max(np.abs(out.real - expect)) < 1e-4
    # what remains is the periodic truncation of e^{-r^2} at |x| = 4 (e^{-16} ~ 1e-7)
    assert np.max(np.abs(out.real - expect)) < 1e-5


def test_norm_and_normalize_reference(oracle_mod):
    """python/test/test.py:123-147; sum of e^{-2 r^2} dV = (pi/2)^{3/2} to machine precision."""
    from gp5_amd import synth

    shape, left, right = (50, 50, 50), (-5.0,) * 3, (5.0,) * 3
    geo = oracle_mod.Geometry(shape, left, right, 1)
    psi = np.asfortranarray(np.exp(-synth.r2_grid(shape, left, right))[..., None].astype(np.complex128))
    assert abs(oracle_mod.norm(psi, geo) - (np.pi / 2) ** 1.5) < 1e-14
    p1 = oracle_mod.normalize(psi, geo, 0, 1.0)
    assert abs(oracle_mod.norm(p1, geo) - 1.0) < 1e-14
    np.testing.assert_allclose(p1 / psi, np.sqrt(1 / (np.pi / 2) ** 1.5), rtol=1e-14)


def test_harmonic_ground_state_energy(oracle_mod):
    """psi = pi^{-3/4} e^{-r^2/2} is the omega = 1 trap ground state: H psi = 1.5 psi and E = 1.5."""
    from gp5_amd import synth

    shape, left, right = (48, 48, 48), (-9.0,) * 3, (9.0,) * 3
    geo = oracle_mod.Geometry(shape, left, right, 1)
    psi = np.pi ** -0.75 * np.exp(-synth.r2_grid(shape, left, right) / 2)
    psi = np.asfortranarray(psi[..., None].astype(np.complex128))
    V = synth.harmonic_potential(shape, left, right, [(1.0, 1.0, 1.0)])
    F = oracle_mod.Functional(kind=oracle_mod.OR_GP, V=V)
    Hpsi = oracle_mod.apply(F, psi, geo)
    assert np.max(np.abs(Hpsi - 1.5 * psi)) < 1e-12
    assert abs(oracle_mod.energy(F, psi, geo) - 1.5) < 1e-12


def _plane_wave_setup(shape=(8, 8, 8), m=(1, 2, 0)):
    from gp5_amd import synth

    left, right = (-2.0,) * 3, (2.0,) * 3
    x, y, z = synth.axes(shape, left, right)
    k = [2 * np.pi * m[d] / 4.0 for d in range(3)]
    f = 0.7 * np.exp(1j * (k[0] * x[:, None, None] + k[1] * y[None, :, None] + k[2] * z[None, None, :]))
    lam = 0.5 * sum(kk * kk for kk in k)
    return shape, left, right, np.asfortranarray(f[..., None]), lam


@pytest.mark.parametrize("imag", [True, False])
def test_stepper_amplification_polynomials(oracle_mod, imag):
    """Free plane wave: Euler multiplies by 1 - z, RK4 by 1 - z + z^2/2 - z^3/6 + z^4/24, z = tau lambda."""
    shape, left, right, f, lam = _plane_wave_setup()
    geo = oracle_mod.Geometry(shape, left, right, 1)
    F = oracle_mod.Functional(kind=oracle_mod.OR_GP)
    dt, n = 0.01, 7
    tau = dt if imag else 1j * dt
    z = tau * lam
    e = oracle_mod.advance(F, f, geo, stepper="eulero", dt=dt, imaginary=imag, nsteps=n)
    r = oracle_mod.advance(F, f, geo, stepper="RK4", dt=dt, imaginary=imag, nsteps=n)
    assert rel_l2(e, (1 - z) ** n * f) < 1e-13
    assert rel_l2(r, (1 - z + z ** 2 / 2 - z ** 3 / 6 + z ** 4 / 24) ** n * f) < 1e-13


def test_rk4_stage_renormalisation_plane_wave(oracle_mod):
    """Per-stage renormalisation (src/stepper.cpp:46-70): on a plane wave every stage vector is
    a multiple of the wave, so the 3-D run reduces to a scalar recurrence on its amplitude."""
    shape, left, right, f, lam = _plane_wave_setup()
    geo = oracle_mod.Geometry(shape, left, right, 1)
    F = oracle_mod.Functional(kind=oracle_mod.OR_GP)
    dt, N, vol = 0.02, 3.0, 4.0 ** 3
    for imag in (True, False):
        tau = dt if imag else 1j * dt

        def nrm(a):
            return a * np.sqrt(N / (abs(a) ** 2 * vol))

        a = 0.7 + 0j
        for _ in range(3):
            k1 = lam * a
            p = nrm(a - tau * 0.5 * k1)
            k2 = lam * p
            p = nrm(a - tau * 0.5 * k2)
            k3 = lam * p
            p = nrm(a - tau * k3)
            k4 = lam * p
            a = nrm(a - (tau / 6) * (((k1 + 2 * k2) + 2 * k3) + k4))
        got = oracle_mod.advance(F, f, geo, stepper="RK4", dt=dt, imaginary=imag, nsteps=3, renormalize=[N])
        assert rel_l2(got, a / 0.7 * f) < 1e-13


def _wave(shape, left, right, m, amp):
    from gp5_amd import synth

    x, y, z = synth.axes(shape, left, right)
    L = [right[d] - left[d] for d in range(3)]
    k = [2 * np.pi * m[d] / L[d] for d in range(3)]
    f = amp * np.exp(1j * (k[0] * x[:, None, None] + k[1] * y[None, :, None] + k[2] * z[None, None, :]))
    return f, 0.5 * sum(kk * kk for kk in k)


# Nonlinear branches on plane waves.  With V uniform, |psi_c| is uniform at every
# stage of every step, so H acts on component c as multiplication by a scalar
# h_c(a_0, a_1) and the 3-D run reduces to a recurrence on the amplitudes a_c.
# The scalar h_c comes from the reference's formulas with psi_c = a_c e^{i k_c.x}:
#   gpFunctional, 1 comp. (functional.cpp:13-68):  lam/m + V + g |a|^2
#   gpFunctional, 2 comp. (functional.cpp:71-89):  lam_c/m_c + V_c + g_cc |a_c|^2 + g_01 |a_other|^2
#   unlocked LHY (functional.cpp:125-150):         lam_c + V_c + (r1 + eta r2 + alpha s, eta beta r1 + beta r2
#                                                   + alpha beta^2 s), s = (r1 + beta r2)^{3/2}
# and the energy density is uniform too: e = sum_c (lam_c/m_c + V_c) r_c + e_int(r).
PW_SHAPE, PW_LEFT, PW_RIGHT = (8, 6, 10), (-2.0, -1.5, -2.5), (2.0, 1.5, 2.5)
PW_VOL = 4.0 * 3.0 * 5.0


def _pw_cases():
    # asymmetric coupling input: couplings(0,1) = 0.45 must be used for both components,
    # couplings(1,0) = -0.8 never (functional.cpp:76)
    G2 = np.array([[1.1, 0.45], [-0.8, 0.6]])
    return {
        "gp_g": dict(nc=1, m=[(1, -2, 1)], amp=[0.9], masses=(1.0, 1.0), V=[0.3],
                     ofunc=dict(kind=0, couplings=np.array([[1.7]])),
                     h=lambda lam, r, V: [lam[0] + V[0] + 1.7 * r[0]],
                     e=lambda r: 0.5 * 1.7 * r[0] ** 2),
        "gp_g_mass": dict(nc=1, m=[(2, 1, -1)], amp=[0.6], masses=(2.5, 1.0), V=[0.0],
                          ofunc=dict(kind=0, couplings=np.array([[-0.9]]), masses=(2.5, 1.0)),
                          h=lambda lam, r, V: [lam[0] / 2.5 + V[0] - 0.9 * r[0]],
                          e=lambda r: 0.5 * -0.9 * r[0] ** 2),
        "gp_2c": dict(nc=2, m=[(1, 0, 2), (-1, 1, 0)], amp=[0.8, 0.5], masses=(1.0, 1.7), V=[0.2, -0.1],
                      ofunc=dict(kind=0, couplings=G2, masses=(1.0, 1.7)),
                      h=lambda lam, r, V: [lam[0] + V[0] + 1.1 * r[0] + 0.45 * r[1],
                                           lam[1] / 1.7 + V[1] + 0.6 * r[1] + 0.45 * r[0]],
                      e=lambda r: 0.5 * 1.1 * r[0] ** 2 + 0.5 * 0.6 * r[1] ** 2 + 0.45 * r[0] * r[1]),
        "unlocked": dict(nc=2, m=[(0, 1, -1), (2, 0, 1)], amp=[0.7, 0.9], masses=(1.0, 1.0), V=[0.1, 0.25],
                         ofunc=dict(kind=2, alpha=0.6, beta=0.8, eta=-0.9),
                         h=lambda lam, r, V: [
                             lam[0] + V[0] + r[0] - 0.9 * r[1] + 0.6 * (r[0] + 0.8 * r[1]) ** 1.5,
                             lam[1] + V[1] - 0.9 * 0.8 * r[0] + 0.8 * r[1] + 0.6 * 0.64 * (r[0] + 0.8 * r[1]) ** 1.5],
                         e=None),
        "lhy": dict(nc=1, m=[(1, 1, 1)], amp=[1.2], masses=(1.0, 1.0), V=[0.0],
                    ofunc=dict(kind=1, lhy_a=-2.7, lhy_b=2.2),
                    h=lambda lam, r, V: [lam[0] + V[0] - 2.7 * r[0] + 2.2 * r[0] ** 1.5],
                    e=lambda r: 0.5 * -2.7 * r[0] ** 2 + 0.4 * 2.2 * r[0] ** 2.5),
    }


def _pw_field(case):
    nc = case["nc"]
    f = np.empty(PW_SHAPE + (nc,), dtype=np.complex128, order="F")
    lam = []
    for c in range(nc):
        w, l_c = _wave(PW_SHAPE, PW_LEFT, PW_RIGHT, case["m"][c], 1.0)
        f[..., c] = w
        lam.append(l_c)
    V = np.empty(PW_SHAPE + (nc,), order="F")
    for c in range(nc):
        V[..., c] = case["V"][c]
    return f, lam, V


def _pw_recurrence(case, lam, stepper, tau, nsteps, N=None):
    """Scalar restatement of euleroStepper / RK4Stepper (stepper.cpp:30-70) on the amplitudes."""
    nc = case["nc"]
    hfun = case["h"]
    V = case["V"]

    def H(a):
        h = hfun(lam, [abs(x) ** 2 for x in a], V)
        return [h[c] * a[c] for c in range(nc)]

    def nrm(a):
        if N is None:
            return a
        return [a[c] * np.sqrt(N[c] / (abs(a[c]) ** 2 * PW_VOL)) for c in range(nc)]

    def axpy(a, cf, k):
        return [a[c] - cf * k[c] for c in range(nc)]

    a = [complex(x) for x in case["amp"]]
    for _ in range(nsteps):
        if stepper == "eulero":
            a = nrm(axpy(a, tau, H(a)))
            continue
        k1 = H(a)
        k2 = H(nrm(axpy(a, tau * 0.5, k1)))
        k3 = H(nrm(axpy(a, tau * 0.5, k2)))
        k4 = H(nrm(axpy(a, tau, k3)))
        s = [((k1[c] + 2 * k2[c]) + 2 * k3[c]) + k4[c] for c in range(nc)]
        a = nrm(axpy(a, tau / 6, s))
    return a


@pytest.mark.parametrize("case", ["gp_g", "gp_g_mass", "gp_2c", "unlocked", "lhy"])
def test_nonlinear_apply_plane_wave(oracle_mod, case):
    """H of a plane wave (or of one plane wave per component) is h_c times the wave."""
    cs = _pw_cases()[case]
    f, lam, V = _pw_field(cs)
    a = np.asfortranarray(f * np.array(cs["amp"]))
    geo = oracle_mod.Geometry(PW_SHAPE, PW_LEFT, PW_RIGHT, cs["nc"])
    got = oracle_mod.apply(oracle_mod.Functional(V=V, **cs["ofunc"]), a, geo)
    h = cs["h"](lam, [x * x for x in cs["amp"]], cs["V"])
    for c in range(cs["nc"]):
        assert rel_l2(got[..., c], h[c] * a[..., c]) < 1e-13, c


@pytest.mark.parametrize("case", ["gp_g", "gp_g_mass", "gp_2c", "unlocked", "lhy"])
@pytest.mark.parametrize("stepper,imag,renorm", [("RK4", True, None), ("RK4", False, None), ("RK4", False, [2.0, 0.7]),
                                                  ("eulero", True, None), ("eulero", False, [1.5, 3.0])])
def test_nonlinear_steppers_plane_wave(oracle_mod, case, stepper, imag, renorm):
    """The 3-D oracle run equals the scalar amplitude recurrence on every nonlinear branch.
    Renormalised runs are taken in real time: in imaginary time h_c is real, so a
    renormalised plane wave keeps its amplitude whatever h_c is, and the check would pin
    nothing."""
    cs = _pw_cases()[case]
    nc = cs["nc"]
    f, lam, V = _pw_field(cs)
    a0 = np.asfortranarray(f * np.array(cs["amp"]))
    geo = oracle_mod.Geometry(PW_SHAPE, PW_LEFT, PW_RIGHT, nc)
    dt, nsteps = 0.01, 5
    N = None if renorm is None else renorm[:nc]
    got = oracle_mod.advance(oracle_mod.Functional(V=V, **cs["ofunc"]), a0, geo, stepper=stepper, dt=dt,
                             imaginary=imag, nsteps=nsteps, renormalize=N)
    tau = dt if imag else 1j * dt
    amp = _pw_recurrence(cs, lam, stepper, tau, nsteps, N)
    for c in range(nc):
        assert rel_l2(got[..., c], amp[c] * f[..., c]) < 1e-13, c
    if case == "gp_2c":
        # the transposed coupling entry would give another trajectory: the check can tell them apart
        wrong = dict(cs, h=lambda lam, r, V: [lam[0] + V[0] + 1.1 * r[0] - 0.8 * r[1],
                                               lam[1] / 1.7 + V[1] + 0.6 * r[1] - 0.8 * r[0]])
        amp_w = _pw_recurrence(wrong, lam, stepper, tau, nsteps, N)
        assert rel_l2(got[..., 0], amp_w[0] * f[..., 0]) > 1e-4


@pytest.mark.parametrize("case", ["gp_g", "gp_g_mass", "gp_2c", "lhy"])
def test_energy_plane_wave(oracle_mod, case):
    """E of plane waves = volume * (sum_c (lam_c/m_c + V_c) r_c + e_int(r))."""
    cs = _pw_cases()[case]
    f, lam, V = _pw_field(cs)
    a = np.asfortranarray(f * np.array(cs["amp"]))
    geo = oracle_mod.Geometry(PW_SHAPE, PW_LEFT, PW_RIGHT, cs["nc"])
    r = [x * x for x in cs["amp"]]
    kin = sum((lam[c] / cs["masses"][c] + cs["V"][c]) * r[c] for c in range(cs["nc"]))
    expect = PW_VOL * (kin + cs["e"](r))
    got = oracle_mod.energy(oracle_mod.Functional(V=V, **cs["ofunc"]), a, geo)
    assert got == pytest.approx(expect, rel=1e-13, abs=1e-13)


CASES = {
    "lhy": (1, dict(kind="lhy"), dict(kind=1)),
    "gp_g_V": (1, dict(kind="gp", couplings=[[1.3]]), dict(kind=0, couplings=np.array([[1.3]]))),
    "gp_2c": (2, dict(kind="gp", couplings=[[1.0, 0.4], [0.4, 0.7]], masses=(1.0, 2.0)),
              dict(kind=0, couplings=np.array([[1.0, 0.4], [0.4, 0.7]]), masses=(1.0, 2.0))),
    "unlocked": (2, dict(kind="unlocked", alpha=0.6, beta=0.8, eta=-0.9),
                 dict(kind=2, alpha=0.6, beta=0.8, eta=-0.9)),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("stepper,imag,renorm", [("RK4", True, True), ("RK4", False, False), ("eulero", True, True),
                                                  ("eulero", False, False)])
def test_oracle_matches_numpy_restatement(oracle_mod, case, stepper, imag, renorm):
    from gp5_amd import synth

    nc, npk, ork = CASES[case]
    shape, left, right = (12, 10, 8), (-4.0, -3.5, -3.0), (4.0, 3.5, 3.0)
    r2 = synth.r2_grid(shape, left, right)
    a = np.empty(shape + (nc,), dtype=np.complex128, order="F")
    for c in range(nc):
        a[..., c] = np.exp(-r2 / (2 * (1.2 + 0.2 * c) ** 2)) * (1 + 0.05 * seeded_field(shape, 1, 9 + c)[..., 0])
    V = synth.harmonic_potential(shape, left, right, [(1.0, 0.9, 1.1), (0.6, 0.7, 0.8)][:nc])
    Nn = [2.0, 0.5][:nc] if renorm else None
    ref = np_ref.advance(a, shape, left, right, stepper=stepper, dt=2e-3, imaginary=imag, nsteps=6,
                         renormalize=Nn, V=V, **npk)
    got = oracle_mod.advance(oracle_mod.Functional(V=V, **ork), a, oracle_mod.Geometry(shape, left, right, nc),
                             stepper=stepper, dt=2e-3, imaginary=imag, nsteps=6, renormalize=Nn)
    assert rel_l2(got, ref) < 1e-13


def test_phase_constraint_matches_numpy_restatement(oracle_mod):
    from gp5_amd import synth

    shape, left, right = (12, 10, 8), (-4.0,) * 3, (4.0,) * 3
    x, y, _ = synth.axes(shape, left, right)
    phase = np.asfortranarray(np.broadcast_to(np.arctan2(y[None, :, None], x[:, None, None] + 0.03), shape)[..., None])
    a = np.asfortranarray(np.exp(-synth.r2_grid(shape, left, right) / 3)[..., None].astype(np.complex128))
    ref = np_ref.advance(a, shape, left, right, stepper="RK4", dt=1e-3, imaginary=True, nsteps=4,
                         renormalize=[1.0], phase=phase, kind="lhy")
    got = oracle_mod.advance(oracle_mod.Functional(kind=1), a, oracle_mod.Geometry(shape, left, right, 1),
                             stepper="RK4", dt=1e-3, imaginary=True, nsteps=4, renormalize=[1.0], phase=phase)
    assert rel_l2(got, ref) < 1e-13


def test_seeded_streams_agree(oracle_mod):
    """gp5_amd.synth.uniform and the oracle's C stream give bit-identical inputs."""
    from gp5_amd import synth

    for seed, start in [(1, 0), (2, 12345), (7, 2 ** 33)]:
        np.testing.assert_array_equal(synth.uniform(seed, 1000, start), oracle_mod.uniform(seed, 1000, start))
