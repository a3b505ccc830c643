"""Pins the batched Gaussian-fit oracle: its analytic gradient against central
finite differences of its loss, for every parameter column, and planted
mistakes that the check must catch."""
import numpy as np
import pytest

from oracle import beamlet_oracle as B
from oracle import gauss_fit_oracle as GF

LAM = 2.51e-12


def small_instance(seed=0, n=3, shape=(40, 48), finite_radius=True):
    rng = np.random.default_rng(seed)
    r1 = B.grid_coords((1e-6, 1e-6), shape)
    p = np.zeros((n, GF.NPARAM))
    p[:, 0] = rng.uniform(-10e-6, 10e-6, n)
    p[:, 1] = rng.uniform(-8e-6, 8e-6, n)
    p[:, 2:4] = rng.uniform(-1e-7, 1e-7, (n, 2))
    p[:, 4] = rng.uniform(0, 1e-12, n)
    amp = rng.uniform(0.3, 1.0, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    p[:, 5], p[:, 6] = amp.real, amp.imag
    p[:, 7:9] = rng.uniform(4e-6, 9e-6, (n, 2))
    p[:, 9:11] = np.inf
    if finite_radius:
        p[0, 9:11] = (80.0, -120.0)
    p[:, 11] = LAM
    p[:, 12] = rng.uniform(-np.pi / 2, np.pi / 2, n)
    tgt = p.copy()
    tgt[:, 0:2] += rng.normal(0, 1e-6, (n, 2))
    tgt[:, 7:9] *= 1 + rng.normal(0, 0.05, (n, 2))
    target = np.abs(GF.field(tgt, r1)) ** 2
    return p, r1, target


def _colwise_err(a, b):
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
    return np.abs(a - b).max(axis=0) / scale


def test_grad_matches_finite_differences_every_column():
    p, r1, target = small_instance()
    g = GF.grad(p, r1, target)
    fd = GF.fd_grad(p, r1, target)
    err = _colwise_err(g[:, :13], fd[:, :13])
    assert np.all(err < 1e-6), err
    assert np.all(g[:, 13:] == 0)
    # every column is exercised (non-trivial gradient)
    assert np.all(np.abs(fd[:, :13]).max(axis=0) > 0)


def test_grad_with_infinite_radii_has_zero_radius_column():
    p, r1, target = small_instance(seed=1, finite_radius=False)
    g = GF.grad(p, r1, target)
    assert np.all(g[:, 9:11] == 0)
    fd = GF.fd_grad(p, r1, target, cols=[0, 1, 7, 8, 12])
    err = _colwise_err(g[:, [0, 1, 7, 8, 12]], fd[:, [0, 1, 7, 8, 12]])
    assert np.all(err < 1e-6), err


@pytest.mark.parametrize("mistake, cols", [("centre_sign", [0, 1]), ("drop_tilt_in_centre", [0, 1]),
                                           ("swap_axes", [7])])
def test_planted_mistakes_are_caught(mistake, cols):
    p, r1, target = small_instance(seed=2)
    # a large tilt makes the centre's tilt term matter
    p[:, 2:4] *= 20
    bad = GF.grad(p, r1, target, _mistake=mistake)
    fd = GF.fd_grad(p, r1, target, cols=cols)
    assert np.max(_colwise_err(bad[:, cols], fd[:, cols])) > 1e-3


def test_loss_and_grad_vanish_at_the_target():
    p, r1, _ = small_instance(seed=3)
    target = np.abs(GF.field(p, r1)) ** 2
    assert GF.loss(p, r1, target) == 0.0
    assert np.all(GF.grad(p, r1, target) == 0.0)


def test_empty_instance():
    r1 = B.grid_coords((1e-6, 1e-6), (5, 7))
    assert np.all(GF.field(np.zeros((0, 16)), r1) == 0)
    assert GF.grad(np.zeros((0, 16)), r1, np.ones(35)).shape == (0, 16)


def test_magnitude_bounds_every_gradient_entry():
    p, r1, target = small_instance(seed=4)
    g = GF.grad(p, r1, target)
    m = GF.grad(p, r1, target, magnitude=True)
    assert np.all(m[:, :13] >= np.abs(g[:, :13]) * (1 - 1e-12))
    assert np.all(m[:, 13:] == 0)
