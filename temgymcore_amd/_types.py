"""Coordinate and shape tuples of the public API (reference __init__.py:16-125)."""
from __future__ import annotations

from typing import NamedTuple, Union

import numpy as np
import torch

PositiveFloat = float
NonNegativeFloat = float
Radians = float
Degrees = float


class ShapeYX(NamedTuple):
    """Image shape as (height, width) in pixels."""

    y: int
    x: int


class ScaleYX(NamedTuple):
    """Pixel size in metres per pixel along (y, x)."""

    y: float
    x: float


class CoordXY(NamedTuple):
    """One continuous coordinate (x, y) in metres."""

    x: float
    y: float

    def to_coords(self) -> "CoordsXY":
        return CoordsXY(x=np.array((self.x,)), y=np.array((self.y,)))


class CoordsXY(NamedTuple):
    """Continuous coordinates (x, y) in metres; arrays of any broadcastable shape."""

    x: Union[np.ndarray, torch.Tensor]
    y: Union[np.ndarray, torch.Tensor]


class PixelYX(NamedTuple):
    """One 0-based pixel coordinate (y, x)."""

    y: Union[int, float]
    x: Union[int, float]

    def to_pixels(self) -> "PixelsYX":
        return PixelsYX(x=np.array((self.x,)), y=np.array((self.y,)))


class PixelsYX(NamedTuple):
    """0-based pixel coordinates (y, x); integer or float arrays."""

    y: Union[np.ndarray, torch.Tensor]
    x: Union[np.ndarray, torch.Tensor]
