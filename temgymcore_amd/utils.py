"""The helpers of the reference's ``temgym_core.utils`` (reference
utils.py:1-325) under their names, for code that imports them from there.

The sampling and ray-matrix helpers are those of ``source`` and
``transfer``.  ``inplace_sum`` is the binning's masked, bounds-checked
scatter-add (numpy arrays, or tensors through torch's indexed accumulate).
The rest are the small numerical helpers the reference's examples and tests
use: Fresnel propagation by its transfer function, a two-propagation lens
imaging solution, a circular aperture mask and phase referencing.
"""
from __future__ import annotations

import numpy as np
import torch

from .source import concentric_rings, fibonacci_spiral, multi_cumsum_inplace, random_coords
from .transfer import custom_jacobian_matrix

__all__ = [
    "FresnelPropagator", "concentric_rings", "custom_jacobian_matrix", "fibonacci_spiral",
    "fresnel_lens_imaging_solution", "inplace_sum", "make_aperture", "multi_cumsum_inplace", "random_coords",
    "try_ravel", "try_reshape", "zero_phase",
]


def inplace_sum(px_y, px_x, mask, frame, buffer) -> None:
    """buffer[py_i, px_i] += frame[i] for every i with mask[i] and (py_i, px_i)
    inside buffer's [h, w]; buffer changes in place (reference
    utils.py:83-114).  numpy arrays, or torch tensors (on buffer's device);
    an integer buffer adds the values cast to its dtype."""
    h, w = buffer.shape
    if isinstance(buffer, torch.Tensor):
        dev = buffer.device
        py = torch.as_tensor(px_y, device=dev).reshape(-1).to(torch.int64)
        px = torch.as_tensor(px_x, device=dev).reshape(-1).to(torch.int64)
        keep = torch.as_tensor(mask, device=dev).reshape(-1).to(torch.bool)
        keep = keep & (py >= 0) & (py < h) & (px >= 0) & (px < w)
        vals = torch.as_tensor(frame, device=dev).reshape(-1)[keep].to(buffer.dtype)
        buffer.index_put_((py[keep], px[keep]), vals, accumulate=True)
        return None
    py = np.ravel(np.asarray(px_y)).astype(np.int64)
    px = np.ravel(np.asarray(px_x)).astype(np.int64)
    keep = np.ravel(np.asarray(mask)).astype(bool) & (py >= 0) & (py < h) & (px >= 0) & (px < w)
    np.add.at(buffer, (py[keep], px[keep]), np.ravel(np.asarray(frame))[keep].astype(buffer.dtype))
    return None


def try_ravel(val):
    """``val.ravel()`` where val has it, else val (reference utils.py:208-224)."""
    ravel = getattr(val, "ravel", None)
    return val if ravel is None else ravel()


def try_reshape(val, maybe_has_shape):
    """val reshaped to ``maybe_has_shape.shape`` where both allow it, else val
    (reference utils.py:227-245)."""
    shape = getattr(maybe_has_shape, "shape", None)
    reshape = getattr(val, "reshape", None)
    return val if shape is None or reshape is None else reshape(shape)


def FresnelPropagator(u1, L, wavelength, z, xp=np):
    """The field u1 ([M, N], side length L, square pixels L / M) propagated
    over z in the Fresnel approximation: multiplied in the (unshifted) FFT
    domain by exp(-i pi wavelength z (fx^2 + fy^2)) (reference
    utils.py:248-265).  xp: numpy or a numpy-like module."""
    m, n = u1.shape
    pitch = L / m
    fy = xp.fft.fftfreq(m, d=pitch)[:, None]
    fx = xp.fft.fftfreq(n, d=pitch)[None, :]
    transfer = xp.exp(-1j * xp.pi * wavelength * z * (fx**2 + fy**2))
    return xp.fft.ifft2(transfer * xp.fft.fft2(u1))


def fresnel_lens_imaging_solution(E0, Y, X, ps, lambda0, z1, f, z2):
    """E0 propagated over z1, through a thin lens of focal length f (phase
    exp(-i k (X^2 + Y^2) / (2 f)), X, Y the grid in metres) and over z2
    (reference utils.py:268-275); ps is the pixel size."""
    k = 2 * np.pi / lambda0
    side = E0.shape[0] * ps
    at_lens = FresnelPropagator(E0, side, lambda0, z1) * np.exp(-1j * k / (2 * f) * (X**2 + Y**2))
    return FresnelPropagator(at_lens, side, lambda0, z2)


def zero_phase(u, idx_x, idx_y):
    """u rotated in place so that u[idx_x, idx_y] has phase 0; returns u
    (reference utils.py:278-282)."""
    u *= np.exp(-1j * np.angle(u[idx_x, idx_y]))
    return u


def make_aperture(X, Y, aperture_ratio=0.1):
    """Circular mask X^2 + Y^2 < r^2 with r = aperture_ratio * max |X|
    (reference utils.py:285-294)."""
    radius = aperture_ratio * np.max(np.abs(X))
    return X**2 + Y**2 < radius**2
