"""Pins the 4D-STEM oracle (float64) with closed forms, reductions to the
pinned reference components, FD convergence and bilinear identities."""
import math

import numpy as np
import pytest

from oracle import ray_oracle as R
from oracle import stem_oracle as S

GEOM = dict(scan_shape=(5, 6), n_side=8, semi_conv=0.5e-3, defocus=2e-7, pivot=1e4, offsets=(0.0, 0.0, 0.0, 0.0))


def _params(rng, cams=(0.1, 0.55, 1.0)):
    A = rng.normal(0, 1e-3, 16)
    return np.concatenate([A, [1e-10, 12.0], cams])


def test_descanner4_reduces_to_reference_descanner():
    rng = np.random.default_rng(0)
    e = rng.normal(0, 1e-2, 12)
    A = [[e[0], e[1], 0.0, 0.0], [e[2], e[3], 0.0, 0.0], [e[4], e[5], 0.0, 0.0], [e[6], e[7], 0.0, 0.0]]
    r = R.make_ray(rng.normal(size=50), rng.normal(size=50), rng.normal(size=50), rng.normal(size=50))
    a = S.descanner4(r, 0.3, -0.7, 0.01, 0.02, A, (e[8], e[9], e[10], e[11]))
    b = R.descanner(r, 0.3, -0.7, 0.01, 0.02, e)
    for k in ("x", "y", "dx", "dy"):
        np.testing.assert_array_equal(a[k], b[k])


def test_scan_positions_step_vector_and_centering():
    spx, spy = S.scan_positions((64, 64), 1e-10, 12.0)
    # zero mean over the symmetric grid, and one column step = (cos t, -sin t) * step
    assert abs(spx.mean()) < 1e-22 and abs(spy.mean()) < 1e-22
    t = math.radians(12.0)
    np.testing.assert_allclose([spx[1] - spx[0], spy[1] - spy[0]], [1e-10 * math.cos(t), -1e-10 * math.sin(t)],
                               rtol=1e-12)


def test_forward_affine_closed_form():
    rng = np.random.default_rng(1)
    p = _params(rng)
    x, y = S.forward(GEOM, p)
    A, step, rot, cams = S.unpack(p)
    spx, spy = S.scan_positions(GEOM["scan_shape"], step, rot)
    t = S.slopes(GEOM["n_side"], GEOM["semi_conv"]).astype(np.float64)
    ty, tx = np.meshgrid(t, t, indexing="ij")
    kap, d = GEOM["pivot"], GEOM["defocus"]
    v = np.stack([spx, spy, kap * spx, kap * spy])  # [4, S]
    for c, L in enumerate(cams):
        bx = kap * spx * d + A[0] @ v + A[2] @ v * L
        by = kap * spy * d + A[1] @ v + A[3] @ v * L
        np.testing.assert_allclose(x[c], tx.reshape(1, -1) * (d + L) + bx[:, None], rtol=0, atol=1e-18)
        np.testing.assert_allclose(y[c], ty.reshape(1, -1) * (d + L) + by[:, None], rtol=0, atol=1e-18)


def test_fd_gradient_converges_and_matches_exact_camera_derivative():
    rng = np.random.default_rng(2)
    p = _params(rng)
    q = p * (1 + 1e-3 * rng.normal(size=p.shape))
    tx, ty = S.forward(GEOM, q)
    targets = np.stack([tx, ty], axis=-1)
    g1 = S.fd_grad(GEOM, p, targets, rel=1e-2)
    g2 = S.fd_grad(GEOM, p, targets, rel=3e-3)
    steps = S.fd_steps(p)
    steps[:16] = 1e-1
    g3 = S.fd_grad(GEOM, p, targets, steps=steps)
    for blk in (slice(0, 16), slice(16, 17), slice(17, 18), slice(18, None)):
        scale = np.abs(g1[blk]).max()
        np.testing.assert_allclose(g1[blk], g2[blk], rtol=0, atol=1e-5 * scale)
        np.testing.assert_allclose(g1[blk], g3[blk], rtol=0, atol=1e-5 * scale)
    # the map enters linearly: g[A_0k] = v_k sum 2 r_x, so A_02 / A_00 = pivot
    assert g1[2] == pytest.approx(GEOM["pivot"] * g1[0], rel=1e-5)
    assert g1[7] == pytest.approx(GEOM["pivot"] * g1[5], rel=1e-5)
    # exact d loss / d L_c = sum 2 r (dx at the descanner output)
    x, y = S.forward(GEOM, p)
    A, step, rot, cams = S.unpack(p)
    spx, spy = S.scan_positions(GEOM["scan_shape"], step, rot)
    t = S.slopes(GEOM["n_side"], GEOM["semi_conv"]).astype(np.float64)
    tyy, txx = np.meshgrid(t, t, indexing="ij")
    kap = GEOM["pivot"]
    v = np.stack([spx, spy, kap * spx, kap * spy])
    ddx = txx.reshape(1, -1) + (A[2] @ v)[:, None]
    ddy = tyy.reshape(1, -1) + (A[3] @ v)[:, None]
    for c in range(len(cams)):
        exact = np.sum(2 * (x[c] - targets[c, ..., 0]) * ddx) + np.sum(2 * (y[c] - targets[c, ..., 1]) * ddy)
        assert g1[18 + c] == pytest.approx(exact, rel=1e-6)


def test_planted_sign_error_in_descanner_is_caught():
    # a descanner that adds +spx instead of -spx breaks the closed form
    r = R.make_ray(0.0, 0.0, 0.0, 0.0)
    good = S.descanner4(r, 1e-9, 0.0, 0.0, 0.0, np.zeros((4, 4)), (0, 0, 0, 0))["x"]
    assert good == pytest.approx(-1e-9)


def test_bilinear_identities():
    Tinv = R.metres_to_pixels_matrix((0.0, 0.0), (1e-10, 1e-10), (16, 16))
    yy, xx = np.meshgrid(np.arange(16.0), np.arange(16.0), indexing="ij")
    img = 0.3 * yy - 0.7 * xx + 2.0  # linear image
    rng = np.random.default_rng(3)
    x = rng.uniform(-6e-10, 6e-10, 500)
    y = rng.uniform(-6e-10, 6e-10, 500)
    fy, fx = R.transform(y, x, Tinv)
    np.testing.assert_allclose(S.bilinear(img, Tinv, x, y), 0.3 * fy - 0.7 * fx + 2.0, atol=1e-12)
    # pixel centres return pixel values
    T = R.pixels_to_metres_matrix((0.0, 0.0), (1e-10, 1e-10), (16, 16))
    my, mx = R.transform(np.array([3.0]), np.array([11.0]), T)
    assert S.bilinear(img, Tinv, mx, my)[0] == pytest.approx(img[3, 11], abs=1e-12)
    # outside the image: zero
    assert S.bilinear(img, Tinv, np.array([1e-6]), np.array([0.0]))[0] == 0.0


def _bp_setup(seed=9):
    geom = dict(scan_shape=(6, 5), n_side=8, semi_conv=0.5e-3, defocus=200e-9, pivot=1e4,
                offsets=(0.0, 0.0, 0.0, 0.0), sample_shape=(40, 44))
    rng = np.random.default_rng(seed)
    p = np.concatenate([rng.normal(0, 1e-3, 16), [1e-10, 12.0], [0.5, 0.8]])
    det_Tinv = R.metres_to_pixels_matrix((0.0, 0.0), (2e-5, 2e-5), (64, 60))
    img_Tinv = R.metres_to_pixels_matrix((0.0, 0.0), (2e-11, 2e-11), geom["sample_shape"])
    return geom, p, det_Tinv, img_Tinv, rng


def _deposit(geom, p, img, det_Tinv, img_Tinv, shape, cam=0):
    """4D deposit of bilinear sample values into [S, H, W], from the pinned
    forward, pixel transform and bilinear lookup."""
    xd, yd = S.forward(geom, p)
    xs, ys = S.sample_plane(geom, p)
    v = S.bilinear(img, img_Tinv, xs, ys)
    out = np.zeros((xd.shape[1],) + shape)
    for s in range(xd.shape[1]):
        py, px = R.metres_to_pixels(xd[cam, s], yd[cam, s], det_Tinv)
        ok = (py >= 0) & (py < shape[0]) & (px >= 0) & (px < shape[1])
        np.add.at(out[s], (py[ok], px[ok]), v[s][ok])
    return out


def test_backproject_is_the_transpose_of_the_deposit():
    """<deposit(I), D> == <I, backproject(D)> for random I, D and both
    cameras; a transposed sample-pixel transform breaks it."""
    geom, p, det_Tinv, img_Tinv, rng = _bp_setup()
    for cam in (0, 1):
        img = rng.normal(size=geom["sample_shape"])
        D = rng.normal(size=(30, 64, 60))
        lhs = np.sum(_deposit(geom, p, img, det_Tinv, img_Tinv, (64, 60), cam) * D)
        B, Wt = S.backproject(geom, p, D, det_Tinv, img_Tinv, cam)
        assert abs(np.sum(img * B) - lhs) <= 1e-12 * np.sum(np.abs(img) * np.abs(B)) + 1e-300
        assert np.all(Wt >= 0) and Wt.sum() > 0
    swapped = img_Tinv[[1, 0, 2]]
    B_bad, _ = S.backproject(geom, p, D, det_Tinv, swapped, 1)
    assert abs(np.sum(img * B_bad) - lhs) > 1e-3 * abs(lhs)


def test_backproject_weights_sum_to_ray_count():
    """Bilinear weights of a ray sum to 1 when all four neighbours are on the
    image, so the weight image sums to the number of rays that land on both
    the detector and the image interior."""
    geom, p, det_Tinv, img_Tinv, rng = _bp_setup(10)
    D = np.ones((30, 64, 60))
    B, Wt = S.backproject(geom, p, D, det_Tinv, img_Tinv, 0)
    np.testing.assert_allclose(B, Wt, rtol=1e-15, atol=0)
    n_rays = 30 * 64
    assert abs(Wt.sum() - n_rays) < 1e-9 * n_rays


def test_grad_abs_scale_bounds_the_gradient_and_is_tight_without_cancellation():
    """|sum_j d(r_j^2)/dp| <= sum_j |d(r_j^2)/dp| for every parameter, with
    equality when every term has one sign: a single scan position whose
    x residuals are all +1e-9 m and y residuals 0 makes d(r^2)/dA[0][0] =
    2 * 1e-9 * spx for every ray."""
    rng = np.random.default_rng(4)
    p = _params(rng)
    t = np.stack(S.forward(GEOM, _params(rng)), -1)
    g = S.fd_grad(GEOM, p, t)
    scale = S.grad_abs_scale(GEOM, p, t)
    assert np.all(scale >= np.abs(g) * (1 - 1e-9))
    assert np.any(scale > 2 * np.abs(g))  # the terms of some sums cancel
    idx = [7]
    x, y = S.forward(GEOM, p, idx)
    one = np.stack([x - 1e-9, y], -1)
    g1 = S.fd_grad(GEOM, p, one, scan_idx=idx)
    s1 = S.grad_abs_scale(GEOM, p, one, scan_idx=idx)
    assert s1[0] == pytest.approx(abs(g1[0]), rel=1e-6)
    spx, _ = S.scan_positions(GEOM["scan_shape"], p[16], p[17])
    n = len(p[18:]) * GEOM["n_side"] ** 2
    assert s1[0] == pytest.approx(2e-9 * abs(spx[7]) * n, rel=1e-6)
