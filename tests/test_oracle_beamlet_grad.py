"""Pins the finite-difference beamlet gradients of beamlet_grad_oracle."""
import numpy as np

from oracle import beamlet_grad_oracle as BG
from oracle import beamlet_oracle as B
from oracle import gauss_fit_oracle as GF


def _params(n, rng, lam=2e-12, w=5e-8, spread=1e-7, tilt=1e-5, curved=True):
    p = np.zeros((n, 16))
    p[:, 0:2] = rng.uniform(-spread, spread, (n, 2))
    p[:, 2:4] = rng.uniform(-tilt, tilt, (n, 2))
    p[:, 4] = rng.uniform(0, 1e-10, n)
    a = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    p[:, 5], p[:, 6] = a.real, a.imag
    p[:, 7:9] = w * rng.uniform(0.7, 1.3, (n, 2))
    p[:, 9:11] = np.inf
    if curved:
        p[0, 9:11] = (0.8, -1.3)
    p[:, 11] = lam
    p[:, 12] = rng.uniform(-1, 1, n)
    return p


def test_input_mode_matches_fit_gradient():
    """With g = (4/N)(|u|^2 - T) u the field cotangent is that of the fit's
    intensity loss, whose analytic gradient is pinned by finite differences."""
    rng = np.random.default_rng(0)
    p = _params(3, rng, lam=2.51e-12, w=6e-6, spread=8e-6, tilt=1e-7)
    r = B.grid_coords((1e-6, 1e-6), (30, 34))
    tgt = p.copy()
    tgt[:, 0] += 1e-6
    T = np.abs(GF.field(tgt, r)) ** 2
    u = GF.field(p, r)
    g = (4.0 / r.shape[0]) * (np.abs(u) ** 2 - T) * u
    grad, _, S, _ = BG.grad_fd(p, None, BG.INPUT, r, g)
    ref = GF.grad(p, r, T)
    assert np.all(np.abs(grad[:, :13] - ref[:, :13]) <= 1e-7 * np.maximum(S[:, :13], 1e-300) + 1e-30)


def test_free_space_translation_identity():
    """For A = I (free space) moving the input centre by d moves the field by
    d: du/dx0 = -du/dr_x; the same holds for the offset entry e_x (T[0, 4])."""
    rng = np.random.default_rng(1)
    p = _params(1, rng, curved=False)
    L = 1e-3
    T = np.eye(5)
    T[0, 2] = T[1, 3] = L
    r = B.grid_coords((3e-8, 3e-8), (20, 22))
    g = rng.normal(size=r.shape[0]) + 1j * rng.normal(size=r.shape[0])
    grad, gT, S, _ = BG.grad_fd(p, T[None], BG.PROPAGATED, r, g)
    h = 1e-11
    du_dr = (BG.one_field(p[0], T, BG.PROPAGATED, r + [h, 0]) - BG.one_field(p[0], T, BG.PROPAGATED, r - [h, 0])) / (2 * h)
    ref = np.real(np.sum(np.conj(g) * -du_dr))
    assert abs(grad[0, 0] - ref) <= 1e-6 * S[0, 0]
    assert abs(gT[0, 4] - ref) <= 1e-6 * S[0, 0]


def test_amplitude_columns_closed_form():
    rng = np.random.default_rng(2)
    p = _params(2, rng)
    T = np.eye(5)
    T[0, 2] = T[1, 3] = 2e-3
    r = B.grid_coords((3e-8, 3e-8), (16, 18))
    g = rng.normal(size=r.shape[0]) + 1j * rng.normal(size=r.shape[0])
    grad, _, S, _ = BG.grad_fd(p, np.stack([T, T]), BG.PROPAGATED, r, g)
    for b in range(2):
        u = BG.one_field(p[b], T, BG.PROPAGATED, r)
        a = p[b, 5] + 1j * p[b, 6]
        E = u / a
        assert abs(grad[b, 5] - np.real(np.sum(np.conj(g) * E))) <= 1e-9 * S[b, 5]
        assert abs(grad[b, 6] - np.real(np.sum(np.conj(g) * 1j * E))) <= 1e-9 * S[b, 6]


def test_log_field_derivatives_are_quadratic_in_pixels():
    """Phi is a quadratic in the pixel coordinates for any parameter values,
    so every exact d ln u / dtheta is one too: a least-squares quadratic fit
    of the finite-difference derivatives leaves no residual (this caught a
    step that was too large for weakly coupled parameters)."""
    rng = np.random.default_rng(3)
    p = _params(3, rng, w=5e-8, spread=2e-7)
    p[:, 9:11] = (1.44, 1.63)
    T = np.eye(5)
    T[0, 2] = T[1, 3] = 1e-3
    T[0, 4], T[2, 4] = 2e-8, 1e-5
    H, W = 40, 44
    r = B.grid_coords((3e-8, 3e-8), (H, W))
    py, px = np.divmod(np.arange(H * W), W)
    for b in range(3):
        u0 = BG.one_field(p[b], T.reshape(25), BG.PROPAGATED, r)
        mag = np.abs(u0)
        sig = mag > 1e-4 * mag.max()
        uu, vv = (px[sig] - px[sig].mean()) / 10, (py[sig] - py[sig].mean()) / 10
        M = np.stack([uu * uu, uu * vv, vv * vv, uu, vv, np.ones_like(uu)], 1).astype(np.complex128)
        dirs = [("p", c) for c in range(13) if c not in (5, 6)] + [("T", e) for e in range(20)]
        for kind, c in dirs:
            def f(v, kind=kind, c=c):
                q, TT = p[b].copy(), T.reshape(25).copy()
                (q if kind == "p" else TT)[c] = v
                return BG.one_field(q, TT, BG.PROPAGATED, r)
            x0 = p[b, c] if kind == "p" else T.reshape(25)[c]
            dphi = BG.derivative(f, x0)[sig] / u0[sig]
            coef, *_ = np.linalg.lstsq(M, dphi, rcond=None)
            resid = np.abs(M @ coef - dphi).max() / max(np.abs(dphi).max(), 1e-300)
            assert resid < 1e-6, (b, kind, c, resid)
