"""4D-STEM oracle (numpy float64) -- test infrastructure, see oracle/__init__.py.

Evaluates the shadow-image model ray by ray with the pinned component
restatements of ray_oracle (scanner, propagation, descanner) in float64:

  ray (0, 0, tx, ty) -> scanner(sp, st = pivot*sp) -> propagate(d)
  -> descanner with a 4x4 map -> propagate(L_c)

Scan positions come from the pinned pixel transform (ray_oracle, reference
coordinate_transforms.py:50-101).  The loss gradient is taken by central
finite differences of the float64 loss (steps chosen by ``fd_steps``).

Pinning (tests/test_oracle_stem.py): descanner4 == ray_oracle.descanner when
the tilt columns are zero; scan positions match the reference's rotated
step-vector closed form; FD gradients converge in h and agree with the exact
derivative for parameters that enter linearly; bilinear lookup reproduces
linear images exactly and pixel values at pixel centres; grad_abs_scale
bounds |fd_grad| and equals it (and its closed form 2 r spx N) when every
term has one sign.
"""
from __future__ import annotations

import numpy as np

from . import ray_oracle as R


def slopes(n_side, semi_conv):
    k = np.arange(n_side, dtype=np.float64)
    return ((k - 0.5 * (n_side - 1)) * (2.0 * semi_conv / n_side)).astype(np.float32)


def scan_positions(shape, step, rot_deg):
    T = R.pixels_to_metres_matrix((0.0, 0.0), (step, step), shape, rotation=rot_deg)
    yy, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    my, mx = R.transform(yy.ravel(), xx.ravel(), T)
    return mx, my


def descanner4(r, spx, spy, stx, sty, A, off):
    """4x4 descan map; reduces to ray_oracle.descanner for zero tilt columns."""
    o = r["one"]

    def row(k):
        return spx * A[k][0] + spy * A[k][1] + stx * A[k][2] + sty * A[k][3] + off[k]

    return dict(r, x=r["x"] + (row(0) - spx) * o, y=r["y"] + (row(1) - spy) * o,
                dx=r["dx"] + (row(2) - stx) * o, dy=r["dy"] + (row(3) - sty) * o)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return p[:16].reshape(4, 4), p[16], p[17], p[18:]


def forward(geom, p, scan_idx=None):
    """Detector (x, y) float64 of shape [C, S', J] for scan indices `scan_idx`.

    geom: dict(scan_shape, n_side, semi_conv, defocus, pivot, offsets)."""
    A, step, rot, cams = unpack(p)
    spx, spy = scan_positions(geom["scan_shape"], step, rot)
    if scan_idx is not None:
        spx, spy = spx[scan_idx], spy[scan_idx]
    t = slopes(geom["n_side"], geom["semi_conv"]).astype(np.float64)
    ty, tx = np.meshgrid(t, t, indexing="ij")
    tx, ty = tx.reshape(1, -1), ty.reshape(1, -1)
    spx, spy = spx.reshape(-1, 1), spy.reshape(-1, 1)
    stx, sty = geom["pivot"] * spx, geom["pivot"] * spy
    d = geom["defocus"]
    xs, ys = [], []
    for L in cams:
        r = R.make_ray(0.0, 0.0, tx, ty)
        r = R.scanner(r, spx, spy, stx, sty)
        r = R.propagate(r, d - r["z"])
        r = descanner4(r, spx, spy, stx, sty, A, geom["offsets"])
        r = R.propagate(r, (d + L) - r["z"])
        xs.append(np.broadcast_to(r["x"], (spx.shape[0], tx.shape[1])))
        ys.append(np.broadcast_to(r["y"], (spx.shape[0], tx.shape[1])))
    return np.stack(xs), np.stack(ys)


def sample_plane(geom, p):
    """(x, y) at the sample plane, float64 [S, J]."""
    _, step, rot, _ = unpack(p)
    spx, spy = scan_positions(geom["scan_shape"], step, rot)
    t = slopes(geom["n_side"], geom["semi_conv"]).astype(np.float64)
    ty, tx = np.meshgrid(t, t, indexing="ij")
    spx, spy = spx.reshape(-1, 1), spy.reshape(-1, 1)
    r = R.make_ray(0.0, 0.0, tx.reshape(1, -1), ty.reshape(1, -1))
    r = R.scanner(r, spx, spy, geom["pivot"] * spx, geom["pivot"] * spy)
    r = R.propagate(r, geom["defocus"])
    return r["x"], r["y"]


def bilinear(img, Tinv, x, y):
    """Bilinear lookup at metric (x, y); neighbours outside contribute zero."""
    fy, fx = R.transform(y, x, Tinv)
    y0, x0 = np.floor(fy), np.floor(fx)
    wy, wx = fy - y0, fx - x0
    h, w = img.shape
    out = np.zeros(np.broadcast(fy, fx).shape)
    for a in (0, 1):
        for b in (0, 1):
            yy, xx = (y0 + a).astype(np.int64), (x0 + b).astype(np.int64)
            ok = (yy >= 0) & (yy < h) & (xx >= 0) & (xx < w)
            wgt = (wy if a else 1 - wy) * (wx if b else 1 - wx)
            out += np.where(ok, wgt * img[np.clip(yy, 0, h - 1), np.clip(xx, 0, w - 1)], 0.0)
    return out


def loss(geom, p, targets, scan_idx=None):
    """sum over (c, s, j) of |detector - target|^2 (float64); targets [C, S', J, 2]."""
    x, y = forward(geom, p, scan_idx)
    t = np.asarray(targets, dtype=np.float64)
    return float(np.sum((x - t[..., 0]) ** 2) + np.sum((y - t[..., 1]) ** 2))


def fd_steps(p, rel=1e-2):
    """Central-difference steps.  The loss is exactly quadratic in the descan
    map, the scan step and the camera lengths, so central differences are
    exact for any h there; rotation enters through sin/cos and its O(h^2)
    error is (2e-3 rad)^2/6 ~ 1e-6 relative at rel = 1e-2.  Steps much smaller
    than these (e.g. 1e-4 relative) put the loss change within ~1e-4 of the
    f64 rounding floor of the ~1e-4 m detector coordinates.  The map uses an
    absolute step of 1e-2 (its entries are ~1e-3)."""
    p = np.asarray(p, dtype=np.float64)
    h = rel * np.where(p != 0, np.abs(p), 1.0)
    h[:16] = 1e-2
    return h


def fd_grad(geom, p, targets, scan_idx=None, rel=1e-2, steps=None):
    """Central finite differences of `loss` with one Richardson step:
    g = (4 D(h/2) - D(h)) / 3 removes the O(h^2) term (exact for the
    parameters the loss is quadratic in; O(h^4) for the rotation)."""
    p = np.asarray(p, dtype=np.float64)
    steps = fd_steps(p, rel) if steps is None else np.asarray(steps, dtype=np.float64)
    g = np.zeros_like(p)

    def central(i, h):
        a, b = p.copy(), p.copy()
        a[i] += h
        b[i] -= h
        return (loss(geom, a, targets, scan_idx) - loss(geom, b, targets, scan_idx)) / (2 * h)

    for i in range(p.size):
        h = steps[i]
        g[i] = (4.0 * central(i, 0.5 * h) - central(i, h)) / 3.0
    return g


def grad_abs_scale(geom, p, targets, scan_idx=None, rel=1e-2):
    """sum over the loss terms of |d(r^2)/dp_k| = 2 sum |r| |dr/dp_k| per
    parameter (r the detector residuals, dr/dp_k by central differences of
    the residual fields).  Not a value under test: the scale of one f32
    rounding per term of the gradient sums, which sets how far an f32
    evaluation may sit from the float64 gradient when the sum cancels."""
    p = np.asarray(p, dtype=np.float64)
    t = np.asarray(targets, dtype=np.float64)
    x, y = forward(geom, p, scan_idx)
    rx, ry = x - t[..., 0], y - t[..., 1]
    h = fd_steps(p, rel)
    out = np.zeros_like(p)
    for i in range(p.size):
        a, b = p.copy(), p.copy()
        a[i] += h[i]
        b[i] -= h[i]
        xa, ya = forward(geom, a, scan_idx)
        xb, yb = forward(geom, b, scan_idx)
        out[i] = 2.0 * (np.abs(rx * (xa - xb)).sum() + np.abs(ry * (ya - yb)).sum()) / (2.0 * h[i])
    return out


def backproject(geom, p, data4d, det_Tinv, img_Tinv, cam=0):
    """Back-projection of data4d [S, H, W] onto the sample image (float64):
    every ray whose detector pixel (round half even, int32 cast, in range)
    lies on the detector adds data4d[s, pixel] times its bilinear weights at
    the sample-plane position (same weights and bounds as ``bilinear``).
    Returns (image, weight).  det_Tinv / img_Tinv: metres -> pixels matrices.
    The adjoint of the forward's 4D deposit of ``bilinear`` values."""
    xd, yd = forward(geom, p)
    xd, yd = xd[cam], yd[cam]
    H, W = data4d.shape[1:]
    fy, fx = R.transform(yd, xd, det_Tinv)
    py, px = R.round_to_i32(fy), R.round_to_i32(fx)
    valid = (py >= 0) & (py < H) & (px >= 0) & (px < W)
    s_idx = np.broadcast_to(np.arange(xd.shape[0])[:, None], xd.shape)
    d = np.where(valid, data4d[s_idx, np.clip(py, 0, H - 1), np.clip(px, 0, W - 1)], 0.0)
    xs, ys = sample_plane(geom, p)
    gy, gx = R.transform(ys, xs, img_Tinv)
    y0, x0 = np.floor(gy), np.floor(gx)
    wy, wx = gy - y0, gx - x0
    h, w = geom["sample_shape"]
    inside = valid & (y0 >= -1) & (y0 < h) & (x0 >= -1) & (x0 < w)
    image, weight = np.zeros((h, w)), np.zeros((h, w))
    for a in (0, 1):
        for b in (0, 1):
            yy, xx = (y0 + a).astype(np.int64), (x0 + b).astype(np.int64)
            ok = inside & (yy >= 0) & (yy < h) & (xx >= 0) & (xx < w)
            wgt = (wy if a else 1 - wy) * (wx if b else 1 - wx)
            np.add.at(image, (yy[ok], xx[ok]), (wgt * d)[ok])
            np.add.at(weight, (yy[ok], xx[ok]), wgt[ok])
    return image, weight
