"""Pins the beamlet oracle: the reference's closed-form tests and the
misaligned (shifted / tilted) beamlet against a brute-force Fresnel integral."""
import numpy as np
import pytest

from oracle import beamlet_oracle as B


def test_zr_wz_r_closed_forms():
    # reference test_gaussians.py:167-185
    w0, wl = 1e-3, 500e-9
    zr = B.zR(w0, wl)
    assert zr == pytest.approx(np.pi * w0 * w0 / wl, rel=1e-12)
    assert B.w_z(w0, 0.0, zr) == pytest.approx(w0, rel=1e-12)
    assert B.w_z(w0, zr, zr) == pytest.approx(w0 * np.sqrt(2), rel=1e-12)
    assert B.R(zr, zr) == pytest.approx(2 * zr, rel=1e-12)
    assert B.R(0.0, zr) == np.inf


def test_q_inv_focus_and_rayleigh():
    # reference test_gaussians.py:188-203
    w0, wl = 1e-3, 500e-9
    zr = B.zR(w0, wl)
    assert B.q_inv(0.0, w0, wl) == pytest.approx(1j * wl / (np.pi * w0 * w0), rel=1e-12)
    exp = -1 / B.R(zr, zr) + 1j * wl / (np.pi * B.w_z(w0, zr, zr) ** 2)
    assert B.q_inv(zr, w0, wl) == pytest.approx(exp, rel=1e-12)


def test_gaussian_beam_centre_and_identity_abcd():
    # reference test_gaussians.py:206-226
    q, k = 1j * 1e-6, 2 * np.pi / 500e-9
    assert B.gaussian_beam(0.0, 0.0, q, k) == pytest.approx(1.0)
    assert B.gaussian_beam(1e-3, -2e-3, q, k, -1e-3, 2e-3) == pytest.approx(1.0)
    Q1 = np.array([[2j, 0], [0, 3j]])
    I, Z = np.eye(2), np.zeros((2, 2))
    np.testing.assert_allclose(B.Qinv_ABCD(Q1, I, Z, Z, I), Q1, rtol=1e-12)


def test_aligned_free_space_beamlet():
    # reference test_gaussians.py:229-273: pref * exp(i k/2 r^T Q2 r)
    w0, wl, L = 1e-3, 500e-9, 0.25
    k = 2 * np.pi / wl
    q = 1j * wl / (np.pi * w0 * w0)
    Q1 = np.diag([q, q])
    A, Bm, C, D = np.eye(2), L * np.eye(2), np.zeros((2, 2)), np.eye(2)
    xs = np.array([-1e-4, 0.0, 2e-4])
    ys = np.array([0.0, 1e-4, -2e-4])
    r2 = np.stack(np.meshgrid(xs, ys, indexing="xy"), -1).reshape(-1, 2)
    Q2 = B.Qinv_ABCD(Q1, A, Bm, C, D)
    field = B.beam_field(1.0, 0.0, Q1, Q2, np.zeros(2), np.zeros(2), A, Bm, np.zeros(2), np.zeros(2), k, r2)
    pref = 1.0 / np.sqrt(np.linalg.det(A + Bm @ Q1))
    exp = pref * np.exp(1j * (k / 2) * np.einsum("ni,ij,nj->n", r2, Q2, r2))
    np.testing.assert_allclose(field, exp, rtol=1e-12, atol=1e-14)


def _fresnel_case(shift, tilt, astig=False):
    lam, L, w0 = 1e-6, 1e-3, 20e-6
    k = 2 * np.pi / lam
    waist = np.array([[w0, 0.7 * w0 if astig else w0]])
    radii = np.full((1, 2), np.inf)
    th = np.array([0.4 if astig else 0.0])
    Q1 = B.ray_Q_inv(waist, radii, lam, th)[0]
    xs1 = np.arange(-120e-6, 120e-6, 0.4e-6)
    X, Y = np.meshgrid(xs1, xs1, indexing="xy")
    d = np.stack([X - shift[0], Y - shift[1]], -1)
    u1 = np.exp(1j * (k / 2) * np.einsum("...i,ij,...j->...", d, Q1, d) + 1j * k * (d @ np.asarray(tilt)))
    pts = [(0.0, 0.0), (15e-6, -5e-6), (-20e-6, 12e-6), (30e-6, 25e-6), (5e-6, 35e-6)]
    ref = B.fresnel_direct(u1, xs1, xs1, lam, L, pts)
    A, Bm, C, D = np.eye(2), L * np.eye(2), np.zeros((2, 2)), np.eye(2)
    Q2 = B.Qinv_ABCD(Q1, A, Bm, C, D)
    got = B.beam_field(1.0, 0.0, Q1, Q2, np.asarray(shift, float), np.asarray(tilt, float), A, Bm, np.zeros(2),
                       np.zeros(2), k, np.array(pts))
    return got, ref


@pytest.mark.parametrize("shift, tilt, astig", [
    ((0.0, 0.0), (0.0, 0.0), False),
    ((8e-6, -5e-6), (0.0, 0.0), False),
    ((0.0, 0.0), (4e-3, -2e-3), False),
    ((8e-6, -5e-6), (4e-3, -2e-3), True),
])
def test_misaligned_beamlet_vs_fresnel_integral(shift, tilt, astig):
    got, ref = _fresnel_case(shift, tilt, astig)
    # magnitudes agree; phases agree up to one constant (the reference drops
    # the piston phase of the propagated beam)
    np.testing.assert_allclose(np.abs(got), np.abs(ref), rtol=2e-3, atol=2e-3 * np.abs(ref).max())
    ratio = got / ref
    rel = ratio / ratio[0]
    strong = np.abs(ref) > 1e-2 * np.abs(ref).max()
    np.testing.assert_allclose(np.angle(rel[strong]), 0.0, atol=5e-3)


def test_misaligned_planted_sign_error_is_caught():
    # flipping the tilt sign moves the beam the wrong way: magnitudes disagree
    got, ref = _fresnel_case((8e-6, -5e-6), (4e-3, -2e-3))
    bad, _ = _fresnel_case((8e-6, -5e-6), (-4e-3, 2e-3))
    assert np.max(np.abs(np.abs(bad) - np.abs(ref))) > 0.1 * np.abs(ref).max()


def test_offset_columns_translate_and_tilt():
    # e translates the field, f multiplies it by exp(i k f.r)
    lam, L = 1e-6, 1e-3
    k = 2 * np.pi / lam
    Q1 = B.ray_Q_inv(np.array([[20e-6, 20e-6]]), np.full((1, 2), np.inf), lam, np.zeros(1))[0]
    A, Bm, C, D = np.eye(2), L * np.eye(2), np.zeros((2, 2)), np.eye(2)
    Q2 = B.Qinv_ABCD(Q1, A, Bm, C, D)
    r2 = np.random.default_rng(0).uniform(-3e-5, 3e-5, (50, 2))
    z2 = np.zeros(2)
    base = B.beam_field(1.0, 0.0, Q1, Q2, z2, z2, A, Bm, z2, z2, k, r2)
    e = np.array([3e-6, -1e-6])
    shifted = B.beam_field(1.0, 0.0, Q1, Q2, z2, z2, A, Bm, e, z2, k, r2 + e)
    np.testing.assert_allclose(shifted, base, rtol=1e-10, atol=1e-12)
    f = np.array([2e-3, 1e-3])
    tilted = B.beam_field(1.0, 0.0, Q1, Q2, z2, z2, A, Bm, z2, f, k, r2)
    np.testing.assert_allclose(tilted, base * np.exp(1j * k * (r2 @ f)), rtol=1e-9, atol=1e-12)


def test_sums_are_linear():
    rng = np.random.default_rng(1)
    n = 6
    x, y = rng.uniform(-1e-6, 1e-6, (2, n))
    args = dict(dx=np.zeros(n), dy=np.zeros(n), pathlength=rng.uniform(0, 1e-9, n), amplitude=np.ones(n),
                waist_xy=np.full((n, 2), 5e-8), radii=np.full((n, 2), np.inf), wavelength=np.full(n, 2e-12),
                theta=np.zeros(n))
    r2 = B.grid_coords((2e-8, 2e-8), (16, 16))
    abcd = np.eye(5)
    abcd[0, 2] = abcd[1, 3] = 1e-3
    full = B.make_image(x, y, abcd=abcd, r2=r2, **args)
    half = B.make_image(x[:3], y[:3], abcd=abcd, r2=r2, **{k: v[:3] for k, v in args.items()})
    rest = B.make_image(x[3:], y[3:], abcd=abcd, r2=r2, **{k: v[3:] for k, v in args.items()})
    np.testing.assert_allclose(full, half + rest, rtol=1e-12, atol=1e-12 * np.abs(full).max())
