"""Ray sources generated on the GPU against the oracle's restatement of the
reference generators (utils.py:46-80, :117-205, :297-325; source.py:82-188).

Angles and radii are replayed bit for bit; positions go through the device's
f64 sin/cos, so they are compared within 4 ulp of the radius (f64) or exactly
after rounding the oracle to f32 up to one f32 ulp.  Random discs are checked
by distribution (the reference's sampler is unseeded)."""
import numpy as np
import pytest
import torch

from oracle import ray_oracle as O
from temgymcore_amd import source as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("num", [7, 4096, 250_000])
@pytest.mark.parametrize("kind", ["point", "beam"])
def test_rings_vs_oracle(device, num, kind):
    off = (3e-6, -1e-6)
    if kind == "point":
        src = S.PointSource(z=0.0, semi_conv=5e-4, offset_xy=off)
        ref = O.point_source_array(num, 5e-4, off)
        scale = 5e-4
    else:
        src = S.ParallelBeam(z=0.0, radius=2e-5, offset_xy=off)
        ref = O.parallel_beam_array(num, 2e-5, off)
        scale = 2e-5 + 4e-6
    rays = src.make_rays(num, device=device)
    got = np.stack([rays.x.cpu().numpy(), rays.y.cpu().numpy(), rays.dx.cpu().numpy(), rays.dy.cpu().numpy()], 1)
    assert got.shape[0] == ref.shape[0]
    tol = 4 * np.finfo(np.float64).eps * scale
    assert np.abs(got - ref[:, :4]).max() <= tol
    if kind == "point":
        assert np.all(got[:, 0] == off[0]) and np.all(got[:, 1] == off[1])
    else:
        assert np.all(got[:, 2] == 0.0) and np.all(got[:, 3] == 0.0)


def test_rings_float32(device):
    rays = S.PointSource(z=0.0, semi_conv=5e-4).make_rays(10_000, device=device, dtype=torch.float32)
    ref = O.point_source_array(10_000, 5e-4)
    assert rays.dx.dtype == torch.float32
    got = rays.dx.cpu().numpy()
    ulp = np.spacing(np.abs(ref[:, 2]).astype(np.float32))
    assert np.all(np.abs(got.astype(np.float64) - ref[:, 2].astype(np.float32)) <= ulp)


@pytest.mark.parametrize("n, alpha", [(100_000, 0), (5000, 2), (1, 2)])
def test_fibonacci_vs_oracle(device, n, alpha):
    x, y = S.fibonacci_spiral(n, 0.2e-9, alpha=alpha, device=device)
    rx, ry = O.fibonacci_spiral(n, 0.2e-9, alpha=alpha)
    tol = 4 * np.finfo(np.float64).eps * 0.2e-9
    assert np.abs(x.cpu().numpy() - rx).max() <= tol
    assert np.abs(y.cpu().numpy() - ry).max() <= tol
    # radii are exact: sqrt and the products are correctly rounded
    r_got = np.hypot(x.cpu().numpy(), y.cpu().numpy())
    np.testing.assert_allclose(r_got, np.hypot(rx, ry), rtol=1e-15, atol=0)


def test_random_disc_distribution_and_seed(device):
    n = 1_000_000
    src = S.PointSource(z=0.0, semi_conv=1e-3, offset_xy=(1e-6, 2e-6))
    a = src.make_rays(n, random=True, device=device, seed=1234)
    b = src.make_rays(n, random=True, device=device, seed=1234)
    c = src.make_rays(n, random=True, device=device, seed=1235)
    assert a.dx.numel() == n
    assert torch.equal(a.dx, b.dx) and torch.equal(a.dy, b.dy)
    assert not torch.equal(a.dx, c.dx)
    assert torch.all(a.x == 1e-6) and torch.all(a.y == 2e-6)
    dx, dy = a.dx.cpu().numpy(), a.dy.cpu().numpy()
    u = (dx ** 2 + dy ** 2) / 1e-6  # (r / R)^2 ~ U(0, 1) for area-uniform points
    assert u.max() < 1.0
    us = np.sort(u)
    ks = np.abs(us - (np.arange(n) + 0.5) / n).max()
    assert ks < 2.0 / np.sqrt(n)  # Kolmogorov-Smirnov, far beyond the 1e-3 level
    ang = np.arctan2(dy, dx)
    hist = np.histogram(ang, bins=16, range=(-np.pi, np.pi))[0]
    assert np.abs(hist - n / 16).max() < 5 * np.sqrt(n / 16)
    beam = S.ParallelBeam(z=0.0, radius=1e-5).make_rays(1000, random=True, device=device, seed=7)
    assert torch.all(beam.dx == 0) and torch.all(torch.hypot(beam.x, beam.y) < 1e-5)


def test_device_rays_feed_the_ray_program(device):
    """A device-generated PointSource bundle through a lens: bit-exact to
    the oracle run on the same (downloaded) rays."""
    from temgymcore_amd.components import Detector, Lens
    from temgymcore_amd.run import run_to_end

    src = S.PointSource(z=0.0, semi_conv=5e-4)
    rays = src.make_rays(50_000, device=device)
    model = [src, Lens(z=0.5, focal_length=0.25), Detector(z=1.0, pixel_size=(1e-5, 1e-5), shape=(64, 64))]
    out = run_to_end(rays, model)
    host = [t.cpu().numpy() for t in (rays.x, rays.y, rays.dx, rays.dy)]
    ref = O.run_to_end(O.make_ray(*host, 0.0, 0.0), [("plane", 0.0), ("lens", 0.5, 0.25), ("detector", 1.0)])
    for k in ("x", "y", "dx", "dy", "pathlength"):
        np.testing.assert_array_equal(getattr(out, k).cpu().numpy(), ref[k], err_msg=k)
