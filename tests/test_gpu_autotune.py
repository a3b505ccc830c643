"""Tuned launch choices change speed, not results."""
import numpy as np
import pytest
import torch

from temgymcore_amd import autotune, gauss_fit
from temgymcore_amd.components import Detector, Lens
from temgymcore_amd.gaussian import MODE_INPUT, BeamletSum, beamlet_coefficients
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import plan_run_to_image

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _clean():
    autotune.clear()
    yield
    autotune.clear()


def test_tuned_beamlet_sum_matches_default(device):
    rng = np.random.default_rng(0)
    n, H, W = 3000, 200, 260
    p = np.zeros((n, 16))
    p[:, 0] = rng.uniform(-100e-6, 100e-6, n)
    p[:, 1] = rng.uniform(-80e-6, 80e-6, n)
    p[:, 5] = 1.0
    p[:, 7:9] = rng.uniform(3e-6, 8e-6, (n, 2))
    p[:, 9:11] = np.inf
    p[:, 11] = 2.51e-12
    grid = Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(H, W))
    params = torch.from_numpy(p).to(device)
    coef = beamlet_coefficients(params, grid, MODE_INPUT)
    ref = BeamletSum(n, (H, W), device).run(coef).clone()
    best, times = autotune.tune_beamlet_sum(coef, n, (H, W), candidates=(1, 4, 16, 64))
    assert best in times and set(times) == {1, 4, 16, 64}
    assert autotune.get_choice(autotune.BEAMLET_SPLITS, (n, H, W)) == best
    got = BeamletSum(n, (H, W), device).run(coef)
    scale = ref.abs().max().item()
    assert (got - ref).abs().max().item() <= 1e-6 * scale


def test_tuned_run_to_image_is_bit_exact(device):
    rng = np.random.default_rng(1)
    n = 300_000
    x, y = (torch.from_numpy(rng.uniform(-0.1, 0.1, n).astype(np.float32)).to(device) for _ in range(2))
    dx, dy = (torch.from_numpy(rng.uniform(-0.05, 0.05, n).astype(np.float32)).to(device) for _ in range(2))
    ray = Ray(x=x, y=y, dx=dx, dy=dy, z=0.0, pathlength=0.0)
    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]
    plan = plan_run_to_image(ray, model)
    plan.launch()
    ref = plan.acc.clone()
    best, times = autotune.tune_run_to_image(plan)
    assert best in times
    plan.acc.zero_()
    plan.launch()
    assert torch.equal(plan.acc, ref)


def test_tuned_fit_slicing_matches_default(device):
    rng = np.random.default_rng(2)
    counts, sides = [700, 40], [150, 130]
    grids = [Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(s, s)) for s in sides]
    p = gauss_fit.workload_params(np.array(counts), np.array(sides), seed=3)
    params = torch.from_numpy(p).to(device)
    base = gauss_fit.GaussianFitBatch(grids, counts)
    target = base.intensity(torch.from_numpy(gauss_fit.perturb_params(p)).to(device))
    l0, g0 = base.loss_and_grad(params, target)
    best, times, tuned, shape, shape_times = autotune.tune_gauss_fit(grids, counts, params, target,
                                                                     candidates=(64, 256, 1024))
    assert best in times and shape in shape_times and set(shape_times) == set(autotune.TILE_SHAPES)
    assert int(tuned.plan.tile_shape) == shape
    l1, g1 = tuned.loss_and_grad(params, target)
    torch.testing.assert_close(l1, l0, rtol=1e-5, atol=0)
    scale = g0.abs().amax(dim=0, keepdim=True)
    assert ((g1 - g0).abs() <= 1e-4 * scale + 1e-300).all()


def _rotated_gaussians(rng, n):
    p = np.zeros((n, 16))
    p[:, 0] = rng.uniform(-100e-6, 100e-6, n)
    p[:, 1] = rng.uniform(-80e-6, 80e-6, n)
    p[:, 5] = 1.0
    p[:, 7:9] = rng.uniform(3e-6, 8e-6, (n, 2))
    p[:, 9:11] = np.inf
    p[:, 11] = 2.51e-12
    p[:, 12] = rng.uniform(-1.5, 1.5, n)  # rotated astigmatism: not separable, the VALU recurrence runs
    return p


def test_tile_shapes_give_the_default_field(device):
    """Every tile shape of the recurrence tile sum sums the same pixels of the
    same beamlets in the same order: the 16-pixel-run shapes start their runs
    at the same pixels (bit-exact), the 8-pixel runs restart the recurrence
    more often (rounding only)."""
    rng = np.random.default_rng(5)
    n, H, W = 3000, 200, 260
    grid = Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(H, W))
    params = torch.from_numpy(_rotated_gaussians(rng, n)).to(device)
    coef = beamlet_coefficients(params, grid, MODE_INPUT)
    ref = BeamletSum(n, (H, W), device).run(coef).clone()
    scale = ref.abs().max().item()
    for shape, (_, _, run) in autotune.TILE_SHAPES.items():
        autotune.set_choice(autotune.TILE_SHAPE, (n, H, W), shape)
        got = BeamletSum(n, (H, W), device).run(coef)
        if run == 16:
            assert torch.equal(got, ref), shape
        else:
            assert (got - ref).abs().max().item() <= 2e-6 * scale, shape
    best, times = autotune.tune_beamlet_tile_shape(coef, n, (H, W))
    assert best in times and set(times) == set(autotune.TILE_SHAPES)


def test_fit_tile_shapes_give_the_default_loss_and_grad(device):
    rng = np.random.default_rng(6)
    counts, sides = [300, 25], [170, 90]
    grids = [Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(s, s)) for s in sides]
    p = gauss_fit.workload_params(np.array(counts), np.array(sides), seed=4)
    params = torch.from_numpy(p).to(device)
    base = gauss_fit.GaussianFitBatch(grids, counts)
    target = base.intensity(torch.from_numpy(gauss_fit.perturb_params(p)).to(device))
    f0 = base.field(params).clone()
    l0, g0 = base.loss_and_grad(params, target)
    key = (len(grids), sum(counts), sum(s * s for s in sides), 1)
    for shape, (_, _, run) in autotune.TILE_SHAPES.items():
        autotune.set_choice(autotune.TILE_SHAPE, key, shape)
        b = gauss_fit.GaussianFitBatch(grids, counts)
        assert int(b.plan.tile_shape) == shape
        f1 = b.field(params)
        l1, g1 = b.loss_and_grad(params, target)
        if run == 16:
            assert torch.equal(f1, f0) and torch.equal(l1, l0) and torch.equal(g1, g0), shape
        else:
            assert (f1 - f0).abs().max().item() <= 2e-6 * f0.abs().max().item(), shape
            torch.testing.assert_close(l1, l0, rtol=1e-5, atol=0)
            scale = g0.abs().amax(dim=0, keepdim=True)
            assert ((g1 - g0).abs() <= 1e-4 * scale + 1e-300).all(), shape
