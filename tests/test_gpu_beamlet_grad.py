"""GPU parity of the beamlet adjoint (tg_beamlet_backward) against
finite-difference derivatives of the complex128 oracle field.

Tolerance: every gradient entry |g - ref| <= 1e-4 S, S = sum over pixels of
|conj(cotangent) du_b/dtheta| (the size of the terms the entry sums; f32
fields and run sums), and per column <= 1e-2 of the column's largest entry
where that entry is not itself a cancellation remainder.
"""
import numpy as np
import pytest
import torch

from oracle import beamlet_grad_oracle as BG
from oracle import beamlet_oracle as BO
from temgymcore_amd import gauss_fit
from temgymcore_amd.components import Detector
from temgymcore_amd.gaussian import (MODE_INPUT, MODE_PROPAGATED, GaussianRay, beamlet_backward, beamlet_field,
                                     make_gaussian_image, pack_params)

pytestmark = pytest.mark.gpu


def _rays(n, rng, lam=2e-12, w=5e-8, spread=2e-7, tilt=1e-5):
    p = np.zeros((n, 16))
    p[:, 0:2] = rng.uniform(-spread, spread, (n, 2))
    p[:, 2:4] = rng.uniform(-tilt, tilt, (n, 2))
    p[:, 4] = rng.uniform(0, 1e-10, n)
    a = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    p[:, 5], p[:, 6] = a.real, a.imag
    p[:, 7:9] = w * rng.uniform(0.7, 1.3, (n, 2))
    p[:, 9:11] = rng.uniform(0.5, 2.0, (n, 2))
    p[::3, 9:11] = np.inf
    p[:, 11] = lam
    p[:, 12] = rng.uniform(-1, 1, n)
    return p


def _check(got, ref, S, tol=1e-4, col_tol=1e-2):
    err = np.abs(got - ref)
    bad = err > tol * S + 1e-300
    assert not bad.any(), f"entries {np.argwhere(bad)[:8]}: err/S {(err / np.maximum(S, 1e-300))[bad][:8]}"
    scale = np.abs(ref).max(axis=0)
    live = scale > 1e-2 * S.max(axis=0)
    colbad = live & (err.max(axis=0) > col_tol * scale)
    assert not colbad.any(), f"columns {np.nonzero(colbad)[0]}"


def _cotangent(rng, shape, field=None):
    g = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    if field is not None:  # a smooth part correlated with the field: less cancellation
        g = g + 3.0 * field / np.abs(field).max()
    return g.astype(np.complex64)


def _abcd(kind):
    T = np.eye(5)
    if kind == "free":
        T[0, 2] = T[1, 3] = 1e-3
    elif kind == "lens":  # drift 0.5 mm, lens f = 2 mm, drift 1 mm
        def drift(d):
            M = np.eye(5)
            M[0, 2] = M[1, 3] = d
            return M
        lens = np.eye(5)
        lens[2, 0] = lens[3, 1] = -1 / 2e-3
        T = drift(1e-3) @ lens @ drift(0.5e-3)
    elif kind == "offsets":
        T[0, 2] = T[1, 3] = 1e-3
        T[0, 4], T[1, 4], T[2, 4], T[3, 4] = 3e-8, -2e-8, 2e-5, -1e-5
    return T


@pytest.mark.parametrize("kind", ["free", "lens", "offsets"])
def test_propagated_backward_vs_oracle(device, kind):
    rng = np.random.default_rng(21)
    n = 7
    p = _rays(n, rng)
    T = _abcd(kind)
    abcd = np.repeat(T[None], n, 0)
    abcd[1:, 0, 2] *= rng.uniform(0.9, 1.1, n - 1)  # per-beamlet matrices
    abcd[1:, 1, 3] = abcd[1:, 0, 2]
    grid = Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(72, 80))
    r = BO.grid_coords((3e-8, 3e-8), (72, 80))
    params = torch.from_numpy(p).to(device)
    A = torch.from_numpy(abcd).to(device).contiguous()
    u = beamlet_field(params, grid, MODE_PROPAGATED, A, 25).cpu().numpy()
    g = _cotangent(rng, (72, 80), u)
    gp, ga = beamlet_backward(params, grid, MODE_PROPAGATED, torch.from_numpy(g).to(device), A, 25, want_abcd=True)
    ref, refT, S, ST = BG.grad_fd(p, abcd, BG.PROPAGATED, r, g.reshape(-1).astype(np.complex128))
    _check(gp.cpu().numpy()[:, :13], ref[:, :13], S[:, :13])
    assert np.all(gp.cpu().numpy()[:, 13:] == 0)
    _check(ga.cpu().numpy()[:, :20], refT[:, :20], ST[:, :20])
    assert np.all(ga.cpu().numpy()[:, 20:] == 0)


def test_input_backward_vs_oracle(device):
    rng = np.random.default_rng(22)
    p = _rays(9, rng, lam=2.51e-12, w=6e-6, spread=20e-6, tilt=1e-6)
    grid = Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(64, 70))
    r = BO.grid_coords((1e-6, 1e-6), (64, 70))
    params = torch.from_numpy(p).to(device)
    u = beamlet_field(params, grid, MODE_INPUT).cpu().numpy()
    g = _cotangent(rng, (64, 70), u)
    gp, ga = beamlet_backward(params, grid, MODE_INPUT, torch.from_numpy(g).to(device))
    assert ga is None
    ref, _, S, _ = BG.grad_fd(p, None, BG.INPUT, r, g.reshape(-1).astype(np.complex128))
    _check(gp.cpu().numpy()[:, :13], ref[:, :13], S[:, :13])


def test_input_backward_matches_batched_fit(device):
    """The dual-number adjoint and the fit's closed-form adjoint agree when
    given the fit's own cotangent (4/N)(|u|^2 - T) u."""
    rng = np.random.default_rng(23)
    H, W = 90, 77
    grid = Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(H, W))
    p = _rays(12, rng, lam=2.51e-12, w=8e-6, spread=25e-6, tilt=1e-4)
    params = torch.from_numpy(p).to(device)
    batch = gauss_fit.GaussianFitBatch([grid], [12])
    tgt = p.copy()
    tgt[:, 1] += 1.5e-6
    target = batch.intensity(torch.from_numpy(tgt).to(device))
    _, g_fit = batch.loss_and_grad(params, target)
    u = batch.field(params).to(torch.complex128)
    cot = ((4.0 / (H * W)) * (u.abs() ** 2 - target.double()) * u).reshape(H, W)
    g_bw, _ = beamlet_backward(params, grid, MODE_INPUT, cot)
    a, b = g_bw.cpu().numpy()[:, :13], g_fit.cpu().numpy()[:, :13]
    scale = np.abs(b).max(axis=0)
    assert np.all(np.abs(a - b).max(axis=0) <= 1e-3 * scale + 1e-300)


def test_shared_abcd_gradient_is_the_sum(device):
    rng = np.random.default_rng(24)
    n = 5
    p = _rays(n, rng)
    T = _abcd("lens")
    grid = Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(48, 40))
    params = torch.from_numpy(p).to(device)
    g = torch.from_numpy(_cotangent(rng, (48, 40))).to(device)
    shared = torch.from_numpy(T[None]).to(device)
    per = torch.from_numpy(np.repeat(T[None], n, 0)).to(device).contiguous()
    g1, a1 = beamlet_backward(params, grid, MODE_PROPAGATED, g, shared, 0, want_abcd=True)
    g2, a2 = beamlet_backward(params, grid, MODE_PROPAGATED, g, per, 25, want_abcd=True)
    assert torch.equal(g1, g2) and torch.equal(a1, a2)


def test_autograd_through_make_gaussian_image(device):
    rng = np.random.default_rng(25)
    n = 6
    p = _rays(n, rng)
    x = torch.from_numpy(p[:, 0].copy()).to(device).requires_grad_(True)
    waist = torch.from_numpy(p[:, 7:9].copy()).to(device).requires_grad_(True)
    T = torch.from_numpy(_abcd("free")).to(device).requires_grad_(True)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    rays = GaussianRay(x=x, y=t(p[:, 1]), dx=t(p[:, 2]), dy=t(p[:, 3]), z=t(np.zeros(n)), pathlength=t(p[:, 4]),
                       _one=t(np.ones(n)), amplitude=t(p[:, 5] + 1j * p[:, 6]), waist_xy=waist,
                       radii_of_curv=t(p[:, 9:11]), wavelength=t(p[:, 11]), theta=t(p[:, 12]))
    grid = Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(56, 60))
    img = make_gaussian_image(rays, [grid], abcd=T)
    # L = sum Re(c u): its cotangent dL/dRe + i dL/dIm is exactly conj(c)
    c = torch.from_numpy(_cotangent(rng, (56, 60), img.detach().cpu().numpy())).to(device)
    loss = (img * c).real.sum()
    loss.backward()
    params = pack_params(rays.to_vector(), device).detach()
    cot = c.conj()
    gp, ga = beamlet_backward(params, grid, MODE_PROPAGATED, cot, T.detach()[None], 0, want_abcd=True)
    torch.testing.assert_close(x.grad, gp[:, 0], rtol=0, atol=0)
    torch.testing.assert_close(waist.grad, gp[:, 7:9], rtol=0, atol=0)
    torch.testing.assert_close(T.grad, ga.sum(0).view(5, 5), rtol=1e-12, atol=0)
    # and against the oracle's finite differences for the beamlet centres
    r = BO.grid_coords((3e-8, 3e-8), (56, 60))
    ref, _, S, _ = BG.grad_fd(p, np.repeat(_abcd("free")[None], n, 0), BG.PROPAGATED, r,
                              cot.cpu().numpy().reshape(-1).astype(np.complex128))
    assert np.all(np.abs(x.grad.cpu().numpy() - ref[:, 0]) <= 1e-4 * S[:, 0])


def test_backward_rejects_bad_arguments(device):
    grid = Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(8, 8))
    params = torch.zeros((2, 16), dtype=torch.float64, device=device)
    with pytest.raises(ValueError):
        beamlet_backward(params, grid, MODE_INPUT, torch.zeros((8, 9), dtype=torch.complex64, device=device))
    # propagated mode: one matrix per beamlet (stride 25) must cover every beamlet
    one = torch.eye(5, dtype=torch.float64, device=device)[None]
    cot = torch.zeros((8, 8), dtype=torch.complex64, device=device)
    with pytest.raises(ValueError, match="at least 50 entries"):
        beamlet_backward(params, grid, MODE_PROPAGATED, cot, one, 25)
    with pytest.raises(ValueError, match="at least 50 entries"):
        beamlet_field(params, grid, MODE_PROPAGATED, one, 25)
    with pytest.raises(ValueError, match="abcd_stride"):
        beamlet_field(params, grid, MODE_PROPAGATED, one, 5)
    with pytest.raises(ValueError, match=r"shape \[n, 16\]"):
        beamlet_field(params[:, :15], grid, MODE_INPUT)
