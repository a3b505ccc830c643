"""Grid mixin: pixel<->metre conversion and detector binning (reference grid.py:12-280).

``into_image`` is the GPU replacement of the reference's host round trip
(``np.asarray`` + numba ``inplace_sum``): positions are mapped to pixels in
f64, rounded half-to-even, cast with the int32 semantics of ``astype(int32)``,
bounds-checked and counted into an int32 (or the given accumulator's dtype)
image by the ray-program kernel.  ``run_to_image`` in ``run.py`` fuses the
same binning into the propagation launch.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from . import _lib, _program
from ._types import CoordsXY, PixelsYX
from .coordinate_transforms import apply_transformation, pixels_to_metres_transform

_ACC_DTYPES = {torch.int32: 0, torch.int64: 1, torch.float32: 2, torch.float64: 3}


def _is_device_array(v) -> bool:
    return isinstance(v, torch.Tensor) and v.is_cuda


class Grid:
    """Mixin for components with (z, centre, shape, pixel_size, rotation, flip_y)."""

    @property
    def pixels_to_metres_mat(self) -> np.ndarray:
        return pixels_to_metres_transform(self.centre, self.pixel_size, self.shape, self.flip_y, self.rotation)

    @property
    def metres_to_pixels_mat(self) -> np.ndarray:
        # reference grid.py:50-63 inverts the pixels->metres matrix (LU solve)
        return np.linalg.inv(self.pixels_to_metres_mat)

    @property
    def coords_px(self):
        """Row and column index meshgrids, shape (H, W) (reference grid.py:65-80)."""
        yy, xx = np.meshgrid(np.arange(self.shape[0]), np.arange(self.shape[1]), indexing="ij")
        return yy, xx

    @property
    def coords(self):
        """Pixel centres as (x, y) pairs, row-major, shape (H*W, 2) (reference grid.py:82-100)."""
        yy, xx = self.coords_px
        cx, cy = self.pixels_to_metres((yy.ravel(), xx.ravel()))
        return np.stack((cx, cy), axis=-1).reshape(-1, 2)

    @property
    def coords_1d(self):
        """x of the first row and y of the first column (reference grid.py:102-118)."""
        yy, xx = self.coords_px
        x_coords = self.pixels_to_metres((yy[0, :], xx[0, :]))[0]
        y_coords = self.pixels_to_metres((yy[:, 0], xx[:, 0]))[1]
        return x_coords, y_coords

    def metres_to_pixels(self, coords, cast: bool = True) -> PixelsYX:
        """Metric (x, y) -> pixel (y, x), rounded half-to-even to int32 if `cast`."""
        coords_x, coords_y = coords
        M = self.metres_to_pixels_mat
        if _is_device_array(coords_x) or _is_device_array(coords_y):
            py, px = _device_affine(coords_y, coords_x, M)
            if cast:
                py = _xla_round_i32(py)
                px = _xla_round_i32(px)
            return PixelsYX(y=py, x=px)
        py, px = apply_transformation(np.ravel(coords_y), np.ravel(coords_x), M)
        if cast:
            py = _np_round_i32(py)
            px = _np_round_i32(px)
        py = py.reshape(np.shape(coords_y))
        px = px.reshape(np.shape(coords_x))
        return PixelsYX(y=py, x=px)

    def pixels_to_metres(self, pixels) -> CoordsXY:
        """Pixel (y, x) -> metric (x, y)."""
        pixels_y, pixels_x = pixels
        M = self.pixels_to_metres_mat
        if _is_device_array(pixels_x) or _is_device_array(pixels_y):
            my, mx = _device_affine(pixels_y, pixels_x, M)
            return CoordsXY(x=mx, y=my)
        my, mx = apply_transformation(np.ravel(pixels_y), np.ravel(pixels_x), M)
        my = my.reshape(np.shape(pixels_y))
        mx = mx.reshape(np.shape(pixels_x))
        return CoordsXY(x=mx, y=my)

    def ray_at_grid(self, px_y, px_x, dx=0.0, dy=0.0, z=None):
        """Ray positioned at a (fractional) pixel of this grid (reference grid.py:184-212)."""
        from .ray import Ray

        if z is None:
            z = self.z
        x, y = self.pixels_to_metres((px_y, px_x))
        x = x.item() if np.ndim(x) == 0 else x
        y = y.item() if np.ndim(y) == 0 else y
        return Ray(x=x, y=y, dx=dx, dy=dy, z=z, pathlength=0.0)

    def ray_to_grid(self, ray, cast: bool = False) -> PixelsYX:
        return self.metres_to_pixels((ray.x, ray.y), cast=cast)

    def into_image(self, ray, acc=None):
        """Count rays (or pixel coordinates) into an image of `self.shape`.

        Returns the accumulator: a new int32 device tensor, or `acc` (a device
        tensor of int32/int64/float32/float64, updated in place).
        """
        from .ray import Ray

        device = _lib.require_device()
        acc = _prepare_acc(self.shape, acc, device)
        if isinstance(ray, Ray):
            ys, xs = ray.y, ray.x
            M = self.metres_to_pixels_mat
        else:
            ys, xs = ray
            M = np.eye(3)
        if np.ndim(ys) == 0 and np.ndim(xs) == 0 and not isinstance(ys, torch.Tensor):
            # one (scalar) hit, as the reference's numba loop over a 0-d array
            ys = np.atleast_1d(np.asarray(ys, dtype=np.float64))
            xs = np.atleast_1d(np.asarray(xs, dtype=np.float64))
        def build():
            spec = make_bin_spec(M, self.shape, acc)
            trace = _program.Trace(device)
            sx = trace.lift(xs)
            sy = trace.lift(ys)
            return _program.build_plan(trace, _Pos(sx, sy), (spec, sx, sy))

        # cached by the grid, the position arrays' structure and the
        # accumulator's: a repeated call re-launches the binning program
        _program.plan_cache.launch(("into_image", self, isinstance(ray, Ray), xs, ys, acc), build, acc=acc)
        return acc


@dataclasses.dataclass(frozen=True)
class _Pos:
    """Minimal field container used to bin bare position arrays."""

    x: object
    y: object


def _prepare_acc(shape, acc, device):
    shape = (int(shape[0]), int(shape[1]))
    if acc is None:
        return torch.zeros(shape, dtype=torch.int32, device=device)
    if not isinstance(acc, torch.Tensor) or not acc.is_cuda:
        raise ValueError("acc must be a device tensor (torch CUDA/HIP tensor)")
    if tuple(acc.shape) != shape:
        raise ValueError(f"acc has shape {tuple(acc.shape)}, expected {shape}")
    if acc.dtype not in _ACC_DTYPES:
        raise ValueError(f"acc dtype {acc.dtype} not supported (int32/int64/float32/float64)")
    if not acc.is_contiguous():
        raise ValueError("acc must be contiguous")
    return acc


def make_bin_spec(metres_to_pixels: np.ndarray, shape, acc: torch.Tensor) -> _lib.BinSpec:
    spec = _lib.BinSpec()
    spec.enabled = 1
    spec.acc_dtype = _ACC_DTYPES[acc.dtype]
    spec.x_reg = spec.y_reg = spec.w_reg = -1
    M = np.asarray(metres_to_pixels, dtype=np.float64)
    for k, v in enumerate((M[0, 0], M[0, 1], M[0, 2], M[1, 0], M[1, 1], M[1, 2])):
        spec.m[k] = float(v)
    spec.height, spec.width = int(shape[0]), int(shape[1])
    spec.acc = acc.data_ptr()
    return spec


def _device_affine(a, b, M):
    """T @ [a, b, 1] on device tensors in f64 (same op order as the kernels)."""
    dev = a.device if isinstance(a, torch.Tensor) else b.device
    a = torch.as_tensor(a, device=dev).to(torch.float64)
    b = torch.as_tensor(b, device=dev).to(torch.float64)
    r0 = (a * float(M[0, 0]) + b * float(M[0, 1])) + float(M[0, 2])
    r1 = (a * float(M[1, 0]) + b * float(M[1, 1])) + float(M[1, 2])
    return r0, r1


def _xla_round_i32(v: torch.Tensor) -> torch.Tensor:
    r = torch.round(v)  # half to even
    r = torch.nan_to_num(r, nan=0.0).clamp(-2147483648.0, 2147483647.0)
    return r.to(torch.int32)


def _np_round_i32(v: np.ndarray) -> np.ndarray:
    r = np.rint(np.asarray(v, dtype=np.float64))
    r = np.nan_to_num(r, nan=0.0, posinf=2147483647.0, neginf=-2147483648.0)
    return np.clip(r, -2147483648.0, 2147483647.0).astype(np.int32)
