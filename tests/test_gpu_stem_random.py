"""Seeded random geometries for the 4D-STEM kernels against the float64 oracle.

Each case draws a scan shape (ragged, from 1 x 1), an even ray grid side
(the geometry requires one; sides that do and do not divide the 256-thread
workgroup, up to 300), 1-16 camera lengths,
scan step, rotation, defocus, cone angle, pivot, descan matrix and offsets,
and a random-phase sample image of random shape and pixel size (footprints
inside and beyond the kernel's 32 x 32 LDS tile, and off the image).
Tolerances are those of the fixed-geometry tests: coordinates 1e-6 of
max(|ref|, 1 um), sample values 2e-5, fit loss relative 1e-5 and gradient
entries within 2e-4 of their block's maximum plus one f32 rounding of every
term they sum (2^-24 sum |d(r^2)/dp|, oracle.stem_oracle.grad_abs_scale)
against central differences of the float64 loss on the same f32 targets.
The rounding term matters here: with targets from a 1e-3 perturbation the
residuals are nearly odd in the slopes, so the descan, step and rotation
sums cancel to 1e-3 - 4e-2 of their absolute sums."""
import numpy as np
import pytest
import torch

from oracle import ray_oracle as R
from oracle import stem_oracle as SO
from temgymcore_amd import stem

pytestmark = pytest.mark.gpu


def _geom(m):
    return dict(scan_shape=m.scan_shape, n_side=m.n_side, semi_conv=m.semi_conv, defocus=m.defocus,
                pivot=m.pivot, offsets=tuple(m.descan_offsets))


def _close_pos(got, ref, floor=1e-6):
    bound = 1e-6 * np.maximum(np.abs(ref), floor)
    err = np.abs(got.astype(np.float64) - ref)
    assert np.all(err <= bound), f"max err {err.max():.3e}, worst ratio {(err / bound).max():.3f}"


def _random_model(rng, max_scan, n_side, max_cam):
    n_cam = int(rng.integers(1, max_cam + 1))
    cams = tuple(sorted(float(v) for v in rng.uniform(0.05, 1.5, n_cam)))
    return stem.STEMModel(
        scan_shape=(int(rng.integers(1, max_scan + 1)), int(rng.integers(1, max_scan + 1))),
        scan_step=float(rng.uniform(2e-11, 5e-10)),
        scan_rotation=float(rng.uniform(-180.0, 180.0)),
        n_side=n_side,
        semi_conv=float(rng.uniform(0.1e-3, 2e-3)),
        defocus=float(rng.uniform(20e-9, 600e-9)),
        pivot=float(rng.uniform(-2e4, 2e4)),
        descan=rng.normal(0.0, 1e-3, (4, 4)),
        descan_offsets=tuple(float(v) for v in rng.normal(0.0, 1e-6, 4)),
        camera_lengths=cams,
        sample_shape=(int(rng.integers(64, 700)), int(rng.integers(64, 700))),
        sample_pixel=float(rng.uniform(1e-11, 6e-11)),
    )


@pytest.mark.parametrize("seed", range(12))
def test_random_geometry_forward(device, seed):
    rng = np.random.default_rng(2000 + seed)
    sides = (2, 16, 38, 64, 100, 256, 300)  # every side in turn
    model = _random_model(rng, max_scan=12, n_side=sides[seed % len(sides)], max_cam=16)
    sample = rng.uniform(-np.pi, np.pi, model.sample_shape).astype(np.float32)
    print(f"[random stem forward seed={seed} scan={model.scan_shape} n_side={model.n_side} n_cam={model.n_cam} "
          f"sample={model.sample_shape}]", flush=True)
    res = stem.stem_forward(model, sample=torch.from_numpy(sample).to(device), coords=True, values=True)
    coords = res["coords"].cpu().numpy()
    x, y = SO.forward(_geom(model), model.params())
    _close_pos(coords[..., 0], x)
    _close_pos(coords[..., 1], y)
    xs, ys = SO.sample_plane(_geom(model), model.params())
    Tinv = R.metres_to_pixels_matrix((0.0, 0.0), (model.sample_pixel,) * 2, model.sample_shape)
    ref_v = SO.bilinear(sample.astype(np.float64), Tinv, xs, ys)
    np.testing.assert_allclose(res["values"].cpu().numpy(), ref_v, rtol=0, atol=2e-5)


@pytest.mark.parametrize("seed", range(6))
def test_random_geometry_fit_gradient(device, seed):
    rng = np.random.default_rng(3000 + seed)
    model = _random_model(rng, max_scan=7, n_side=(2, 6, 16, 22)[seed % 4], max_cam=4)
    print(f"[random stem fit seed={seed} scan={model.scan_shape} n_side={model.n_side} n_cam={model.n_cam}]",
          flush=True)
    targets = stem.stem_forward(stem.perturbed(model, seed=seed))["coords"]
    ws = stem.STEMFitWorkspace(model, device)
    out = ws.run(stem.params_tensor(model, device), targets).cpu().numpy()
    t = targets.cpu().numpy()
    ref_loss = SO.loss(_geom(model), model.params(), t)
    ref_g = SO.fd_grad(_geom(model), model.params(), t)
    floor = 2.0 ** -24 * SO.grad_abs_scale(_geom(model), model.params(), t)
    assert out[-1] == pytest.approx(ref_loss, rel=1e-5)
    g = out[:-1]
    for blk in (slice(0, 16), slice(16, 17), slice(17, 18), slice(18, None)):
        bar = 2e-4 * np.abs(ref_g[blk]).max() + floor[blk]
        err = np.abs(g[blk] - ref_g[blk])
        assert np.all(err <= bar), (blk, err / bar)


@pytest.mark.parametrize("seed", range(8))
def test_random_geometry_backproject(device, seed):
    """Back-projection of constant detector images per scan position (a ray on
    a pixel edge cannot matter) onto random sample grids, against the
    oracle's sums: weights within 1e-6 and images within 2e-6 of the largest
    weight, as for the fixed geometries.  40 um detector pixels: every ray
    lands."""
    import dataclasses

    rng = np.random.default_rng(6000 + seed)
    sides = (2, 16, 38, 64, 100, 74, 256, 30)
    model = dataclasses.replace(_random_model(rng, max_scan=9, n_side=sides[seed % len(sides)], max_cam=3),
                                det_pixel=4e-5)
    cam = int(rng.integers(model.n_cam))
    print(f"[random stem backproject seed={seed} scan={model.scan_shape} n_side={model.n_side} cam={cam}/"
          f"{model.n_cam} sample={model.sample_shape}]", flush=True)
    c = rng.uniform(0.5, 2.0, model.n_scan)
    D = np.broadcast_to(c[:, None, None], (model.n_scan,) + tuple(model.det_shape)).astype(np.float32)
    img, w = stem.backproject(model, torch.from_numpy(np.ascontiguousarray(D)).to(device), cam=cam, weights=True)
    geom = dict(_geom(model), sample_shape=tuple(model.sample_shape))
    det_Tinv = R.metres_to_pixels_matrix((0.0, 0.0), (model.det_pixel,) * 2, model.det_shape)
    img_Tinv = R.metres_to_pixels_matrix((0.0, 0.0), (model.sample_pixel,) * 2, model.sample_shape)
    ref_img, ref_w = SO.backproject(geom, model.params(), D.astype(np.float64), det_Tinv, img_Tinv, cam)
    scale = max(ref_w.max(), 1e-300)
    np.testing.assert_allclose(w.cpu().numpy(), ref_w, rtol=0, atol=1e-6 * scale)
    np.testing.assert_allclose(img.cpu().numpy(), ref_img, rtol=0, atol=1e-6 * 2.0 * scale)
