"""Host helpers of the product that mirror reference functions:
temgymcore_amd.transfer (reference transfer.py:6-183, tests/test_transfer.py:7-153)
and the Gaussian-beam closed forms of temgymcore_amd.gaussian (reference
gaussian.py:13-32, :92-110, tests/test_gaussians.py:167-226)."""
import numpy as np
import pytest

from temgymcore_amd import gaussian as G
from temgymcore_amd import transfer as T


def _free_space(d):
    m = np.eye(5)
    m[0, 2] = m[1, 3] = d
    return m


def test_transfer_pt_src_free_space():
    # reference test_transfer.py:7-35
    x0, y0, ang, d = 1.0, -2.0, np.pi / 4, 5.0
    dx0, dy0 = np.cos(ang), np.sin(ang)
    c = T.transfer_rays_pt_src((x0, y0), (np.array([dx0]), np.array([dy0])), _free_space(d))
    np.testing.assert_allclose(c[:, 0], [x0 + d * dx0, y0 + d * dy0, dx0, dy0], rtol=0, atol=1e-12)


@pytest.mark.parametrize("seed", range(4))
def test_transfer_pt_src_random_matrix(seed):
    # reference test_transfer.py:38-63 (seeded here)
    rng = np.random.default_rng(seed)
    Tm = rng.integers(-5, 5, size=(5, 5)).astype(float)
    Tm[4] = [0, 0, 0, 0, 1]
    x0, y0, dx0, dy0 = 0.7, -1.2, 0.3, -0.4
    c = T.transfer_rays_pt_src((x0, y0), (np.array([dx0]), np.array([dy0])), Tm)
    exp = Tm @ np.array([x0, y0, dx0, dy0, 1.0])
    np.testing.assert_allclose(c[:, 0], exp[:4], rtol=0, atol=1e-12)


def test_transfer_pt_src_identity_batch_and_empty():
    # reference test_transfer.py:66-93
    rng = np.random.default_rng(7)
    sx, sy = rng.standard_normal(4), rng.standard_normal(4)
    c = T.transfer_rays_pt_src((0.5, -0.5), (sx, sy), np.eye(5))
    assert c.shape == (4, 4)
    np.testing.assert_array_equal(c[0], 0.5)
    np.testing.assert_array_equal(c[1], -0.5)
    np.testing.assert_array_equal(c[2], sx)
    np.testing.assert_array_equal(c[3], sy)
    empty = T.transfer_rays_pt_src((1.0, 2.0), (np.array([]), np.array([])), np.eye(5))
    assert empty.shape == (4, 0)


def test_negative_distance_transfer():
    # reference test_transfer.py:98-121
    c = T.transfer_rays_pt_src((1.0, 2.0), (np.array([0.1]), np.array([-0.2])), _free_space(-3.0))
    np.testing.assert_allclose(c[:2, 0], [1.0 - 0.3, 2.0 + 0.6], atol=1e-12)


def test_transfer_rays_shape_and_values():
    # reference test_transfer.py:124-136 (shape), plus values: ray n through
    # cumulative matrix m
    rng = np.random.default_rng(3)
    N, M = 10, 7
    assert T.transfer_rays(np.zeros((N, 5)), np.zeros((M, 5, 5))).shape == (N, M, 5)
    rays = np.hstack([rng.standard_normal((N, 4)), np.ones((N, 1))])
    mats = rng.standard_normal((M, 5, 5))
    out = T.transfer_rays(rays, mats)
    cum = T.accumulate_matrices_cumulative(mats)
    for n in range(N):
        for m in range(M):
            np.testing.assert_allclose(out[n, m], cum[m] @ rays[n], rtol=1e-12, atol=1e-12)


def test_accumulate_matrices():
    # reference test_transfer.py:139-153 (diagonal matrices commute, so the
    # order is pinned with non-commuting ones as well)
    A, B, C = np.eye(5), np.eye(5) * 2, np.eye(5) * 3
    for m in (A, B, C):
        m[4, :] = [0, 0, 0, 0, 1]
    np.testing.assert_allclose(T.accumulate_matrices([A, B, C]), C @ B @ A)
    np.testing.assert_allclose(T.accumulate_matrices([C]), C)
    P1, L, P2 = _free_space(0.5), np.eye(5), _free_space(0.25)
    L[2, 0] = L[3, 1] = -1.0
    np.testing.assert_allclose(T.accumulate_matrices([P1, L, P2]), P2 @ L @ P1, rtol=1e-15)
    assert not np.allclose(P2 @ L @ P1, P1 @ L @ P2)


def test_accumulate_matrices_cumulative_order():
    # The reference's loop (transfer.py:178-183) starts from the last matrix
    # and left-multiplies the earlier ones: [M2, M1 M2, M0 M1 M2] for
    # [M0, M1, M2] -- not the running product its docstring describes.  The
    # product keeps the code's order, which transfer_rays depends on.
    rng = np.random.default_rng(9)
    M0, M1, M2 = rng.standard_normal((3, 5, 5))
    cum = T.accumulate_matrices_cumulative(np.stack([M0, M1, M2]))
    assert cum.shape == (3, 5, 5)
    np.testing.assert_allclose(cum[0], M2, rtol=1e-14)
    np.testing.assert_allclose(cum[1], M1 @ M2, rtol=1e-13)
    np.testing.assert_allclose(cum[2], M0 @ M1 @ M2, rtol=1e-13)


def test_ray_matrix_readme_abcd():
    # reference README.md:230-234 through the product's dual-number Jacobian
    from temgymcore_amd.components import Detector, Lens
    from temgymcore_amd.ray import Ray

    J = T.ray_matrix([Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))],
                     Ray(x=0.1, y=0.2, dx=0.3, dy=0.4, z=0.0, pathlength=0.0))
    exp = np.array([[.5, 0, .75, 0, 0], [0, .5, 0, .75, 0], [-1, 0, .5, 0, 0], [0, -1, 0, .5, 0], [0, 0, 0, 0, 1]])
    np.testing.assert_allclose(J, exp, atol=1e-15)
    np.testing.assert_allclose(T.accumulate_matrices([_free_space(0.5), np.array([[1, 0, 0, 0, 0], [0, 1, 0, 0, 0],
                                                                                  [-1, 0, 1, 0, 0], [0, -1, 0, 1, 0],
                                                                                  [0, 0, 0, 0, 1.0]]),
                                                      _free_space(0.5)]), exp, atol=1e-15)


# ---- Gaussian-beam closed forms (reference tests/test_gaussians.py:167-226) --------

def test_zR_wz_R_basic():
    w0, wl = 1e-3, 500e-9
    zr = G.zR(w0, wl)
    assert np.isclose(zr, np.pi * w0 * w0 / wl, rtol=1e-12, atol=0)
    assert np.isclose(G.w_z(w0, 0.0, zr), w0, rtol=1e-12, atol=0)
    assert np.isclose(G.w_z(w0, zr, zr), w0 * np.sqrt(2.0), rtol=1e-12, atol=0)
    assert np.isclose(G.R(zr, zr), 2.0 * zr, rtol=1e-12, atol=0)
    assert G.R(0.0, zr) == np.inf and G.R(5e-11, zr) == np.inf  # the 1e-10 guard
    # R(z) = z + zR^2 / z at a few distances
    for z in (-0.3, 0.01, 2.5):
        assert np.isclose(G.R(z, zr), z + zr * zr / z, rtol=1e-12, atol=0)


def test_q_inv_focus_and_away():
    w0, wl = 1e-3, 500e-9
    zr = G.zR(w0, wl)
    assert np.isclose(G.q_inv(0.0, w0, wl), 1j * wl / (np.pi * w0 * w0), rtol=1e-12, atol=0)
    exp = -(1.0 / G.R(zr, zr)) + 1j * wl / (np.pi * G.w_z(w0, zr, zr) ** 2)
    assert np.isclose(G.q_inv(zr, w0, wl), exp, rtol=1e-12, atol=0)
    # with q(z) = z + i zR: 1/q = 1/R - i wl/(pi w^2), so the reference's
    # q_inv (gaussian.py:99-110) is -1/q(z)
    for z in (0.2, 1.7):
        assert np.isclose(G.q_inv(z, w0, wl), -1.0 / (z + 1j * zr), rtol=1e-12)


def test_gaussian_beam_center_is_one():
    q, k = 1j * 1e-6, 2 * np.pi / 500e-9
    assert np.isclose(G.gaussian_beam(0.0, 0.0, q, k), 1.0 + 0j, rtol=1e-12, atol=0)
    assert np.isclose(G.gaussian_beam(1e-3, -2e-3, q, k, offset_x=-1e-3, offset_y=2e-3), 1.0 + 0j, rtol=1e-12,
                      atol=0)
    # |u| = exp(-r^2/w^2) for q_inv = i wl / (pi w^2)
    w0, wl = 2e-4, 500e-9
    u = G.gaussian_beam(1e-4, 5e-5, G.q_inv(0.0, w0, wl), 2 * np.pi / wl)
    assert np.isclose(abs(u), np.exp(-(1e-4**2 + 5e-5**2) / w0**2), rtol=1e-12)


def test_Qinv_ABCD_identity_and_free_space():
    Q1 = np.array([[2j, 0], [0, 3j]])
    I, Z = np.eye(2), np.zeros((2, 2))
    np.testing.assert_allclose(G.Qinv_ABCD(Q1, I, Z, Z, I), Q1, rtol=1e-12)
    # free space L: 1/q2 = 1/(q1 + L) per axis
    L = 0.3
    q1x, q1y = 1 / 2j, 1 / 3j
    Q2 = G.Qinv_ABCD(Q1, I, L * I, Z, I)
    np.testing.assert_allclose(np.diag(Q2), [1 / (q1x + L), 1 / (q1y + L)], rtol=1e-12)
    np.testing.assert_allclose([Q2[0, 1], Q2[1, 0]], 0, atol=1e-15)
    # a thin lens f: 1/q2 = 1/q1 - 1/f
    f = 0.7
    Q3 = G.Qinv_ABCD(Q1, I, Z, -I / f, I)
    np.testing.assert_allclose(Q3, Q1 - I / f, rtol=1e-12)


def test_decompose_q_inv_round_trip():
    """decompose_Q_inv (reference gaussian.py:35-89) of GaussianRay.Q_inv
    (reference tests/test_gaussians.py:893-950's beams, batched): the larger
    waist first with theta its axis (mod pi), the smaller second; radii are
    1 / Re(1/q), i.e. minus radii_of_curv (the reference's convention), and
    rebuilding with them reproduces |Q_inv| entries' imaginary parts and the
    real parts negated."""
    import torch

    from temgymcore_amd.gaussian import GaussianRay, decompose_Q_inv

    wl = 1.3e-6
    params = [(1e-5, 3e-5, np.inf, np.inf, np.pi / 6), (1.2e-4, 5e-5, 1e-4, 0.1, -np.pi / 4),
              (1e-4, 3e-4, 0.01, np.inf, -np.pi / 3)]
    w = np.array([[a, b] for a, b, *_ in params])
    r = np.array([[c, d] for _, _, c, d, _ in params])
    th = np.array([p[4] for p in params])
    n = len(params)
    gr = GaussianRay(x=torch.zeros(n), y=torch.zeros(n), dx=torch.zeros(n), dy=torch.zeros(n), z=0.0,
                     pathlength=0.0, _one=1.0, amplitude=torch.ones(n), waist_xy=torch.from_numpy(w),
                     radii_of_curv=torch.from_numpy(r), wavelength=wl, theta=torch.from_numpy(th))
    Q = gr.Q_inv
    w1, w2, r1, r2, t = decompose_Q_inv(Q, wl)
    for i, (a, b, c, d, ang) in enumerate(params):
        big, small = (a, b) if a >= b else (b, a)
        rbig, rsmall = (c, d) if a >= b else (d, c)
        axis = ang if a >= b else ang + np.pi / 2
        assert w1[i] == pytest.approx(big, rel=1e-9) and w2[i] == pytest.approx(small, rel=1e-9)
        assert (r1[i] == np.inf if np.isinf(rbig) else r1[i] == pytest.approx(-rbig, rel=1e-9))
        assert (r2[i] == np.inf if np.isinf(rsmall) else r2[i] == pytest.approx(-rsmall, rel=1e-9))
        diff = np.arctan2(np.sin(t[i] - axis), np.cos(t[i] - axis))
        assert min(abs(diff), abs(abs(diff) - np.pi)) < 1e-9
        assert -np.pi / 2 < t[i] <= np.pi / 2
    back = GaussianRay(x=torch.zeros(n), y=torch.zeros(n), dx=torch.zeros(n), dy=torch.zeros(n), z=0.0,
                       pathlength=0.0, _one=1.0, amplitude=torch.ones(n),
                       waist_xy=torch.from_numpy(np.stack([w1, w2], -1)),
                       radii_of_curv=torch.from_numpy(np.stack([r1, r2], -1)), wavelength=wl,
                       theta=torch.from_numpy(t)).Q_inv
    np.testing.assert_allclose(back.imag, Q.imag, rtol=1e-9, atol=1e-9 * np.abs(Q.imag).max())
    np.testing.assert_allclose(back.real, -Q.real, rtol=1e-9, atol=1e-9 * np.abs(Q.real).max())


def test_gaussian_algebra_helpers_and_map_reduce():
    """The reference's 2x2 helpers (gaussian.py:180-222) against explicit
    products, on numpy and torch inputs, and map_reduce (:340-369) as a
    left fold over the leading axis of arrays, tuples and dicts."""
    import torch

    from temgymcore_amd import gaussian as G

    rng = np.random.default_rng(4)
    M1, M2 = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)), rng.normal(size=(2, 2))
    v, w = rng.normal(size=2), rng.normal(size=(5, 2))
    np.testing.assert_allclose(G.matrix_vector_mul(M1, v), M1 @ v)
    np.testing.assert_allclose(G.matrix_matrix_mul(M1, M2), M1 @ M2)
    assert G.matrix_quadratic_mul(v, M1) == pytest.approx(v @ M1 @ v)
    np.testing.assert_allclose(G.matrix_linear_mul(v, M1, w), np.array([v @ M1 @ wn for wn in w]))
    B = rng.normal(size=(3, 2, 2))
    np.testing.assert_allclose(G.matrix_matrix_matrix_mul(B, B, B), B @ B @ np.swapaxes(B, -1, -2))
    t = G.matrix_vector_mul(torch.from_numpy(M2), torch.from_numpy(v))
    assert isinstance(t, torch.Tensor) and np.allclose(t.numpy(), M2 @ v)
    xs = np.arange(1.0, 6.0)
    assert G.map_reduce(lambda x: x * x, lambda a, b: a + b, 0.0, xs) == pytest.approx(55.0)
    assert G.map_reduce(lambda x: x, lambda a, b: a * 10 + b, 0.0, xs) == pytest.approx(12345.0)  # order
    pairs = (np.array([1.0, 2.0]), np.array([3.0, 4.0]))
    assert G.map_reduce(lambda p: p[0] * p[1], lambda a, b: a + b, 0.0, pairs) == pytest.approx(11.0)
    d = {"a": np.array([1.0, 2.0]), "b": np.array([5.0, 7.0])}
    assert G.map_reduce(lambda p: p["b"] - p["a"], lambda a, b: a + b, 0.0, d, batch_size=1) == pytest.approx(9.0)


def test_beamlet_inputs_are_checked_before_any_launch():
    """The beamlet kernels read params [n, 16] f64 and the 5x5 matrices through
    raw pointers: host tensors, other shapes and too few matrices are refused
    on the host."""
    import torch

    from temgymcore_amd.components import Detector

    grid = Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(8, 8))
    with pytest.raises(ValueError, match="device float64"):
        G.beamlet_coefficients(torch.zeros((2, 16), dtype=torch.float64), grid, G.MODE_INPUT)
    with pytest.raises(ValueError, match="device float64"):
        G.beamlet_field(torch.zeros((2, 16), dtype=torch.float32), grid, G.MODE_INPUT)
    with pytest.raises(ValueError, match="device float64"):
        G.beamlet_backward(torch.zeros((2, 16), dtype=torch.float64), grid, G.MODE_INPUT,
                           torch.zeros((8, 8), dtype=torch.complex64))
