"""Gaussian-beamlet oracle (numpy complex128) -- test infrastructure, see oracle/__init__.py.

Restates reference src/temgym_core/gaussian.py op for op:
  zR, w_z, R, q_inv, gaussian_beam          gaussian.py:13-32, :99-110
  Qinv_ABCD                                  gaussian.py:92-96 (left solve)
  q_inv / Q_inv of a GaussianRay             gaussian.py:138-177
  beam_field (the per-beamlet field)         gaussian.py:276-316
  make_image (sum over beamlets)             gaussian.py:225-273, :319-369
  input_image (input-plane Gaussian sum)     gaussian.py:372-427

Pinning (tests/test_oracle_beamlet.py): the reference's closed-form tests
(test_gaussians.py:167-273: zR, w(z), R(z), q_inv, identity Qinv_ABCD, the
aligned free-space beamlet); the misaligned terms (shift, tilt) against a
brute-force Fresnel diffraction integral of the same input Gaussian (up to
the constant phase the reference drops); e/f offset columns against field
translation / plane-wave tilt; linearity of the sums.  input_field /
input_image: the reference's phase-ramp test (phase 0 at the centre, slope
k * tilt; test_gaussians.py:736-786) and rotated-astigmatism test (major
axis at theta; :820-883), the textbook Gaussian beam in its principal frame
(envelope exp(-rho1^2/wx^2 - rho2^2/wy^2), phase k l + k d.tilt -
(k/2)(rho1^2/Rx + rho2^2/Ry)), linearity; planted mistakes (tilt without its
factor 2, the transposed rotation, the pathlength phase sign) each fail.
"""
from __future__ import annotations

import numpy as np


def zR(w0, wl):
    return np.pi * w0**2 / wl


def w_z(w0, z, z_r):
    return w0 * np.sqrt(1 + (z / z_r) ** 2)


def R(z, z_r):
    if abs(z) < 1e-10:
        return np.inf
    return z * (1 + (z_r / z) ** 2)


def q_inv(z, w0, wl):
    z_r = zR(w0, wl)
    if abs(z) < 1e-10:
        return 1j * wl / (np.pi * w0**2)
    return -1.0 / R(z, z_r) + 1j * wl / (np.pi * w_z(w0, z, z_r) ** 2)


def gaussian_beam(x, y, qinv, k, offset_x=0.0, offset_y=0.0):
    return np.exp(1j * k * ((x + offset_x) ** 2 + (y + offset_y) ** 2) / 2 * qinv)


def Qinv_ABCD(Qinv, A, B, C, D):
    """(C + D Qinv) (A + B Qinv)^-1 written as solve(A + B Qinv, C + D Qinv)."""
    return np.linalg.solve(A + B @ Qinv, C + D @ Qinv)


def ray_q_inv(waist_xy, radii, wavelength):
    """Per-axis 1/q of GaussianRay.q_inv (gaussian.py:138-154); arrays [N]."""
    wx, wy = waist_xy[:, 0], waist_xy[:, 1]
    rx, ry = radii[:, 0], radii[:, 1]
    im_x = 1j * wavelength / (np.pi * wx**2)
    im_y = 1j * wavelength / (np.pi * wy**2)
    with np.errstate(divide="ignore"):
        qx = np.where(np.isinf(rx), im_x, -1.0 / rx + im_x)
        qy = np.where(np.isinf(ry), im_y, -1.0 / ry + im_y)
    return qx, qy


def ray_Q_inv(waist_xy, radii, wavelength, theta):
    """R(theta) diag(1/qx, 1/qy) R(theta)^T (gaussian.py:157-177); [N, 2, 2]."""
    qx, qy = ray_q_inv(waist_xy, radii, wavelength)
    n = qx.shape[0]
    Dg = np.zeros((n, 2, 2), dtype=np.complex128)
    Dg[:, 0, 0] = qx
    Dg[:, 1, 1] = qy
    c, s = np.cos(theta), np.sin(theta)
    Rm = np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], -2)
    return np.einsum("nij,njk,npk->nip", Rm, Dg, Rm)


def _safe_inv(M):
    with np.errstate(all="ignore"):
        try:
            out = np.linalg.solve(M, np.eye(2, dtype=M.dtype))
        except np.linalg.LinAlgError:
            out = np.full((2, 2), np.nan, dtype=M.dtype)
    return np.nan_to_num(out, nan=0.0, posinf=0.0, neginf=0.0)


def beam_field(amp, phase_offset, Q1_inv, Q2_inv, r1m, theta1m, A, B, e, f, k, r2):
    """One beamlet at observation points r2 [P, 2] (gaussian.py:276-316)."""
    B_inv = _safe_inv(B)
    Q1 = _safe_inv(Q1_inv)
    r2 = r2 - e
    r2m = A @ r1m + B @ theta1m
    denom = A + B @ Q1_inv
    pref = amp / np.sqrt(np.linalg.det(denom))
    ABinv = A @ B_inv
    phi1 = r1m @ ABinv @ r1m - 2 * np.einsum("i,ij,nj->n", r1m, B_inv, r2)
    AQ1 = A @ Q1
    M = Q1 @ np.linalg.solve(B @ (AQ1 + B), np.eye(2))
    phi2 = r2m @ M @ r2m - 2 * np.einsum("i,ij,nj->n", r2m, M, r2)
    Q2t = np.einsum("ni,ij,nj->n", r2, Q2_inv, r2)
    f_offset = 2 * r2 @ f
    phase = (k / 2) * (Q2t + phi1 - phi2 + f_offset)
    return pref * np.exp(1j * (phase + phase_offset))


def make_image(x, y, dx, dy, pathlength, amplitude, waist_xy, radii, wavelength, theta, abcd, r2):
    """Sum of beamlet fields at r2 [P, 2] (gaussian.py:225-273).

    abcd: [N, 5, 5] (or [5, 5]) per-beamlet ray-transfer matrices.
    """
    n = len(x)
    abcd = np.broadcast_to(np.asarray(abcd, dtype=np.float64), (n, 5, 5))
    Q1 = ray_Q_inv(waist_xy, radii, wavelength, theta)
    k = 2 * np.pi / np.broadcast_to(wavelength, (n,))
    out = np.zeros(r2.shape[0], dtype=np.complex128)
    for i in range(n):
        T = abcd[i]
        A, B, C, D = T[0:2, 0:2], T[0:2, 2:4], T[2:4, 0:2], T[2:4, 2:4]
        e, f = T[0:2, 4], T[2:4, 4]
        Q2 = Qinv_ABCD(Q1[i], A, B, C, D)
        out += beam_field(amplitude[i], k[i] * pathlength[i], Q1[i], Q2, np.array([x[i], y[i]]),
                          np.array([dx[i], dy[i]]), A, B, e, f, k[i], r2)
    return out


def input_field(a, p, q1, r1m, t1m, k, r1):
    """One input-plane Gaussian (gaussian.py:402-407)."""
    d = r1 - r1m
    quad = np.einsum("ni,ij,nj->n", d, q1, d)
    tilt = 2 * d @ t1m
    return a * np.exp(1j * (k / 2) * (quad + tilt)) * np.exp(1j * p)


def input_image(x, y, dx, dy, pathlength, amplitude, waist_xy, radii, wavelength, theta, r1):
    """Sum of input-plane Gaussians (gaussian.py:372-399)."""
    n = len(x)
    Q1 = ray_Q_inv(waist_xy, radii, wavelength, theta)
    k = 2 * np.pi / np.broadcast_to(wavelength, (n,))
    out = np.zeros(r1.shape[0], dtype=np.complex128)
    for i in range(n):
        out += input_field(amplitude[i], k[i] * pathlength[i], Q1[i], np.array([x[i], y[i]]),
                           np.array([dx[i], dy[i]]), k[i], r1)
    return out


def grid_coords(pixel_size, shape, centre=(0.0, 0.0), rotation=0.0, flip_y=False):
    """Pixel centres (x, y) row-major [H*W, 2] of a Grid (reference grid.py:82-100)."""
    from . import ray_oracle as Rr

    T = Rr.pixels_to_metres_matrix(centre, pixel_size, shape, flip_y=flip_y, rotation=rotation)
    yy, xx = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    my, mx = Rr.transform(yy.ravel(), xx.ravel(), T)
    return np.stack([mx, my], axis=-1)


def fresnel_direct(u1, xs1, ys1, lam, L, pts):
    """Brute-force paraxial Fresnel integral of a sampled input field.

    u2(r) = 1/(i lam L) sum u1(r') exp(i k |r - r'|^2 / (2L)) dA'; the constant
    exp(ikL) is omitted, as in the reference's beamlet formula."""
    k = 2 * np.pi / lam
    X, Y = np.meshgrid(xs1, ys1, indexing="xy")
    dA = (xs1[1] - xs1[0]) * (ys1[1] - ys1[0])
    out = []
    for px, py in pts:
        ker = np.exp(1j * k * ((px - X) ** 2 + (py - Y) ** 2) / (2 * L))
        out.append(np.sum(u1 * ker) * dA / (1j * lam * L))
    return np.array(out)
