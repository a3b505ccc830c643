"""3x3 homogeneous pixel<->metre transforms (reference coordinate_transforms.py:7-140).

Host-side float64 numpy: these matrices are per-grid constants handed to the
kernels as f64 coefficients.
"""
from __future__ import annotations

import math

import numpy as np


def _rotate(radians: float) -> np.ndarray:
    # y, x ordering with the y axis flipped (reference coordinate_transforms.py:15-28)
    c, s = math.cos(radians), math.sin(radians)
    return np.array([(c, s, 0.0), (-s, c, 0.0), (0.0, 0.0, 1.0)])


def _rotate_with_deg_to_rad(degrees: float) -> np.ndarray:
    return _rotate(math.pi / 180 * degrees)


def _identity() -> np.ndarray:
    return np.eye(3)


def _scale(pixel_size_yx) -> np.ndarray:
    return np.array([(pixel_size_yx[0], 0.0, 0.0), (0.0, pixel_size_yx[1], 0.0), (0.0, 0.0, 1.0)])


def _shift(centre_yx) -> np.ndarray:
    return np.array([(1.0, 0.0, centre_yx[0]), (0.0, 1.0, centre_yx[1]), (0.0, 0.0, 1.0)])


def _flip_y() -> np.ndarray:
    return np.array([(-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)])


def pixels_to_metres_transform(centre, pixel_size, shape, flip_y=False, rotation=0.0) -> np.ndarray:
    """Transform mapping [y_px, x_px, 1] -> [y_m, x_m, 1].

    Composition order of reference coordinate_transforms.py:50-101:
    flip_cond @ flip_y @ rotate @ shift(centre) @ scale(pixel_size) @ shift(-(shape-1)/2).
    `centre` is consumed in the order it is given (the reference passes the
    (x, y) CoordsXY straight into a (y, x) shift).
    """
    flip = _flip_y() if bool(flip_y) else _identity()
    shape = np.asarray(shape, dtype=np.float64)
    return (
        flip
        @ _flip_y()
        @ _rotate_with_deg_to_rad(float(rotation))
        @ _shift(tuple(float(c) for c in centre))
        @ _scale(tuple(float(p) for p in pixel_size))
        @ _shift(-(shape - 1) / 2.0)
    )


def apply_transformation(y, x, transformation):
    """[y', x'] = T @ [y, x, 1] elementwise (reference coordinate_transforms.py:104-140)."""
    T = np.asarray(transformation, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    y2 = T[0, 0] * y + T[0, 1] * x + T[0, 2]
    x2 = T[1, 0] * y + T[1, 1] * x + T[1, 2]
    return y2, x2
