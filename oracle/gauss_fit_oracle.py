"""Batched Gaussian-decomposition fit oracle (numpy complex128).

Test infrastructure only (see oracle/__init__.py): product code never imports it.

One fit instance is a set of input-plane Gaussians sampled on one detector
grid.  Its field is evaluate_gaussian_input_image (reference
gaussian.py:372-427, restated in beamlet_oracle.input_image) and its loss is
the intensity L2 loss of examples/fit_three_gaussians.ipynb (cell 2):

    L = mean((|u|^2 - T)^2),   T = target intensity.

grad() is the derivative of L with respect to every column of each
Gaussian's parameter row, written directly per pixel.  It is pinned by
central finite differences of loss() in tests/test_oracle_gauss_fit.py
(every column, plus planted mistakes that the check must catch).

Parameter row columns (as tg_beamlet_coefficients):
    0 x, 1 y, 2 dx, 3 dy, 4 pathlength, 5 amp_re, 6 amp_im, 7 waist_x,
    8 waist_y, 9 radius_x, 10 radius_y, 11 wavelength, 12 theta, 13-15 unused
"""
import numpy as np

from oracle import beamlet_oracle as B

NPARAM = 16


def field(params, r1):
    """Complex field [P] of the Gaussians params [n, 16] at points r1 [P, 2]."""
    p = np.asarray(params, dtype=np.float64).reshape(-1, NPARAM)
    if p.shape[0] == 0:
        return np.zeros(r1.shape[0], dtype=np.complex128)
    return B.input_image(p[:, 0], p[:, 1], p[:, 2], p[:, 3], p[:, 4], p[:, 5] + 1j * p[:, 6], p[:, 7:9], p[:, 9:11],
                         p[:, 11], p[:, 12], r1)


def loss(params, r1, target):
    """mean((|u|^2 - T)^2) (fit_three_gaussians.ipynb cell 2)."""
    u = field(params, r1)
    return float(np.mean((np.abs(u) ** 2 - target) ** 2))


def _Q_parts(p):
    """Q1^-1 of each Gaussian and its derivatives with respect to the
    principal-axis 1/q values and theta (gaussian.py:138-177)."""
    wx, wy, rx, ry, lam, th = p[:, 7], p[:, 8], p[:, 9], p[:, 10], p[:, 11], p[:, 12]
    qx, qy = B.ray_q_inv(p[:, 7:9], p[:, 9:11], lam)
    c, s = np.cos(th), np.sin(th)
    Q = B.ray_Q_inv(p[:, 7:9], p[:, 9:11], lam, th)
    dQ_dqx = np.stack([np.stack([c * c, c * s], -1), np.stack([c * s, s * s], -1)], -2)
    dQ_dqy = np.stack([np.stack([s * s, -c * s], -1), np.stack([-c * s, c * c], -1)], -2)
    dd = qx - qy
    dQ_dth = np.stack([np.stack([-2 * c * s * dd, (c * c - s * s) * dd], -1),
                       np.stack([(c * c - s * s) * dd, 2 * c * s * dd], -1)], -2)
    dqx_dwx = -2j * lam / (np.pi * wx**3)
    dqy_dwy = -2j * lam / (np.pi * wy**3)
    with np.errstate(divide="ignore"):
        dqx_drx = np.where(np.isinf(rx), 0.0, 1.0 / rx**2)
        dqy_dry = np.where(np.isinf(ry), 0.0, 1.0 / ry**2)
    return Q, dQ_dqx, dQ_dqy, dQ_dth, dqx_dwx, dqy_dwy, dqx_drx, dqy_dry


def grad(params, r1, target, _mistake=None, magnitude=False, rows=None, u=None):
    """dL/d(params) [n, 16] (columns 13-15 are zero).

    With u = sum_g u_g and W = (4/N)(|u|^2 - T) conj(u),
    dL/dtheta = Re sum_pixels W du_g/dtheta, where
    u_g = a exp(i k pl) exp(i (k/2) [d^T Q d + 2 d.t]), d = r - m.
    magnitude=True returns sum_pixels |W du_g/dtheta| instead: the size of the
    terms each gradient entry sums, i.e. the scale of its rounding error when
    the terms are summed in finite precision.
    rows: only these Gaussians' rows are computed (the others stay zero).
    u: the summed field at r1, if already known (e.g. checked separately on a
    pixel sample); by default it is field(params, r1).
    `_mistake` plants a wrong term for the finite-difference check to catch."""
    p = np.asarray(params, dtype=np.float64).reshape(-1, NPARAM)
    n = p.shape[0]
    out = np.zeros((n, NPARAM))
    if n == 0:
        return out
    u = field(p, r1) if u is None else np.asarray(u, dtype=np.complex128)
    N = r1.shape[0]
    W = (4.0 / N) * (np.abs(u) ** 2 - target) * np.conj(u)
    Q, dQqx, dQqy, dQth, dqx_dwx, dqy_dwy, dqx_drx, dqy_dry = _Q_parts(p)
    for g in (range(n) if rows is None else rows):
        x, y, tx, ty, pl, ar, ai, _, _, _, _, lam, _ = p[g, :13]
        a = ar + 1j * ai
        k = 2 * np.pi / lam
        m = np.array([x, y])
        t = np.array([tx, ty])
        d = r1 - m
        quad = np.einsum("ni,ij,nj->n", d, Q[g], d)
        E = np.exp(1j * k * pl) * np.exp(1j * (k / 2) * (quad + 2 * d @ t))  # u_g / a
        WE = W * E

        def re_sum(dphi):
            if magnitude:
                return float(np.sum(np.abs(WE * a * dphi)))
            return float(np.real(np.sum(WE * a * dphi)))

        Qd = d @ Q[g].T
        tilt_term = t if _mistake != "drop_tilt_in_centre" else np.zeros(2)
        sgn = -1.0 if _mistake != "centre_sign" else 1.0
        out[g, 0] = re_sum(sgn * 1j * k * (Qd[:, 0] + tilt_term[0]))
        out[g, 1] = re_sum(sgn * 1j * k * (Qd[:, 1] + tilt_term[1]))
        out[g, 2] = re_sum(1j * k * d[:, 0])
        out[g, 3] = re_sum(1j * k * d[:, 1])
        out[g, 4] = re_sum(1j * k * np.ones(N))
        out[g, 5] = float(np.sum(np.abs(WE))) if magnitude else float(np.real(np.sum(WE)))
        out[g, 6] = float(np.sum(np.abs(WE))) if magnitude else float(np.real(np.sum(WE * 1j)))

        def dquad(dQ):
            return np.einsum("ni,ij,nj->n", d, dQ, d)

        dQ_wx = dQqx[g] if _mistake != "swap_axes" else dQqy[g]
        out[g, 7] = re_sum(1j * (k / 2) * dquad(dQ_wx * dqx_dwx[g]))
        out[g, 8] = re_sum(1j * (k / 2) * dquad(dQqy[g] * dqy_dwy[g]))
        out[g, 9] = re_sum(1j * (k / 2) * dquad(dQqx[g] * dqx_drx[g]))
        out[g, 10] = re_sum(1j * (k / 2) * dquad(dQqy[g] * dqy_dry[g]))
        # k = 2 pi / lambda; Im Q scales with lambda so (k/2) d^T Im(Q) d does not
        # depend on it: dPhi/dlambda = -(i k / lambda) [pl + d^T Re(Q) d / 2 + d.t]
        quad_re = np.einsum("ni,ij,nj->n", d, Q[g].real, d)
        out[g, 11] = re_sum(-1j * (k / lam) * (pl + 0.5 * quad_re + d @ t))
        out[g, 12] = re_sum(1j * (k / 2) * dquad(dQth[g]))
    return out


def fd_grad(params, r1, target, cols=range(13), rel=1e-5):
    """Central differences of loss() with Richardson extrapolation; the step of
    each column is scaled to the quantity it perturbs."""
    p = np.asarray(params, dtype=np.float64).reshape(-1, NPARAM)
    out = np.zeros_like(p)
    for g in range(p.shape[0]):
        lam = p[g, 11]
        k = 2 * np.pi / lam
        w = min(p[g, 7], p[g, 8])
        for c in cols:
            if c in (0, 1):
                h = rel * w
            elif c in (2, 3):
                h = rel / (k * w)
            elif c == 4:
                h = rel * lam
            elif c in (5, 6):
                h = rel * max(1.0, abs(p[g, 5]) + abs(p[g, 6]))
            elif c in (7, 8):
                h = rel * p[g, c]
            elif c in (9, 10):
                if np.isinf(p[g, c]):
                    continue
                h = rel * abs(p[g, c])
            elif c == 11:
                h = rel * lam
            else:
                h = rel

            def d(hh):
                pp, pm = p.copy(), p.copy()
                pp[g, c] += hh
                pm[g, c] -= hh
                return (loss(pp, r1, target) - loss(pm, r1, target)) / (2 * hh)

            out[g, c] = (4 * d(h / 2) - d(h)) / 3
    return out
