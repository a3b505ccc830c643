"""The reference-named helpers of temgymcore_amd.utils (reference
utils.py:1-325), on the host: the scatter-add against its literal per-entry
loop, Fresnel propagation against the closed form for a Gaussian (which
fixes the transfer function's sign and scale), the lens-imaging solution
against the focal-plane spot of a Gaussian, and the small helpers."""
import numpy as np
import pytest
import torch

from temgymcore_amd import source, transfer, utils


def _loop_sum(py, px, mask, frame, buf):
    h, w = buf.shape
    for i in range(py.size):
        if mask[i] and 0 <= py[i] < h and 0 <= px[i] < w:
            buf[py[i], px[i]] += frame[i]
    return buf


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int64])
def test_inplace_sum_matches_the_per_entry_loop(dtype):
    rng = np.random.default_rng(0)
    n, shape = 2000, (7, 11)
    py, px = rng.integers(-3, 14, n), rng.integers(-3, 14, n)
    mask = rng.uniform(size=n) < 0.7
    frame = (rng.uniform(-2, 2, n) if dtype != np.int64 else rng.integers(-3, 4, n)).astype(dtype)
    ref = _loop_sum(py, px, mask, frame, np.zeros(shape, dtype))
    got = np.zeros(shape, dtype)
    assert utils.inplace_sum(py, px, mask, frame, got) is None
    np.testing.assert_allclose(got, ref, rtol=1e-12 if dtype != np.float32 else 1e-5, atol=1e-5)
    tb = torch.zeros(shape, dtype=torch.from_numpy(np.zeros(1, dtype)).dtype)
    utils.inplace_sum(torch.from_numpy(py), torch.from_numpy(px), torch.from_numpy(mask), torch.from_numpy(frame), tb)
    np.testing.assert_allclose(tb.numpy(), ref, rtol=1e-12 if dtype != np.float32 else 1e-5, atol=1e-5)


def _grid(n, side):
    c = (np.arange(n) - n // 2) * (side / n)
    return np.meshgrid(c, c)  # X, Y


def test_fresnel_propagator_gaussian_closed_form():
    """exp(-r^2 / w0^2) propagated over z is exp(-r^2 / (w0^2 q)) / q with
    q = 1 + i z / zR, zR = pi w0^2 / lambda, for the transfer function
    exp(-i pi lambda z f^2); energy is conserved and propagations compose."""
    n, side, lam, w0 = 256, 2e-4, 1e-9, 6e-6
    X, Y = _grid(n, side)
    u0 = np.exp(-(X**2 + Y**2) / w0**2)
    zr = np.pi * w0**2 / lam
    for z in (0.3 * zr, 1.0 * zr, 2.5 * zr):
        q = 1 + 1j * z / zr
        exact = np.exp(-(X**2 + Y**2) / (w0**2 * q)) / q
        got = utils.FresnelPropagator(u0, side, lam, z)
        assert np.abs(got - exact).max() <= 1e-9
    z1, z2 = 0.4 * zr, 0.9 * zr
    two = utils.FresnelPropagator(utils.FresnelPropagator(u0, side, lam, z1), side, lam, z2)
    np.testing.assert_allclose(two, utils.FresnelPropagator(u0, side, lam, z1 + z2), atol=1e-12)
    assert np.sum(np.abs(two) ** 2) == pytest.approx(np.sum(np.abs(u0) ** 2), rel=1e-12)


def test_lens_imaging_focuses_a_gaussian_to_its_fourier_spot():
    """A Gaussian of waist w0 at a lens (z1 = 0) focuses in the back focal
    plane (z2 = f) to a Gaussian of waist lambda f / (pi w0): |E|^2 second
    moment w^2 / 4 per axis."""
    n, side, lam, w0, f = 512, 4e-4, 1e-9, 4e-5, 0.2
    X, Y = _grid(n, side)
    E0 = np.exp(-(X**2 + Y**2) / w0**2)
    E = utils.fresnel_lens_imaging_solution(E0, Y, X, side / n, lam, 0.0, f, f)
    I = np.abs(E) ** 2
    var_x = float((I * X**2).sum() / I.sum())
    wf = lam * f / (np.pi * w0)
    assert np.sqrt(4 * var_x) == pytest.approx(wf, rel=1e-3)


def test_small_helpers():
    u = np.exp(1j * np.linspace(0, 3, 12)).reshape(3, 4) * 2.0
    out = utils.zero_phase(u, 1, 2)
    assert out is u and abs(np.angle(u[1, 2])) < 1e-15 and np.allclose(np.abs(u), 2.0)
    X, Y = _grid(9, 9.0)
    m = utils.make_aperture(X, Y, aperture_ratio=0.5)
    np.testing.assert_array_equal(m, X**2 + Y**2 < (0.5 * np.abs(X).max()) ** 2)
    a = np.arange(6).reshape(2, 3)
    assert utils.try_ravel(a).shape == (6,) and utils.try_ravel(3.0) == 3.0
    assert utils.try_reshape(np.arange(6), a).shape == (2, 3) and utils.try_reshape(2.0, a) == 2.0
    assert utils.try_reshape(np.arange(6), 1.0).shape == (6,)
    assert utils.concentric_rings is source.concentric_rings and utils.fibonacci_spiral is source.fibonacci_spiral
    assert utils.custom_jacobian_matrix is transfer.custom_jacobian_matrix
    assert utils.random_coords is source.random_coords and utils.multi_cumsum_inplace is source.multi_cumsum_inplace
