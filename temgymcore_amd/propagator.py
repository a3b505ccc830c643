"""Free-space propagators (reference propagator.py:8-116).

``FreeSpaceParaxial`` is the propagator of the batched path: x += dx*d,
y += dy*d, z += d, pathlength += d.  On a vectorised ray a call compiles to a
one-step ray program; inside ``run_to_end`` it is fused with the components.
``FreeSpaceDirCosine`` (direction cosines) serves scalar and host rays.
"""
from __future__ import annotations

import math
import numbers
from typing import NamedTuple, TYPE_CHECKING

import numpy as np
import torch

from . import _program

if TYPE_CHECKING:
    from .ray import Ray


def _propagate_paraxial(ray, distance):
    # reference propagator.py:51-72
    return ray.derive(
        x=ray.x + ray.dx * distance,
        y=ray.y + ray.dy * distance,
        z=ray.z + distance,
        pathlength=ray.pathlength + distance,
    )


class BasePropagator:
    """Abstract propagator: implement ``propagate(ray, distance)``."""

    # a propagator's traced program depends on its type and its attributes
    _plan_key_by_attributes = True

    def __call__(self, ray: "Ray", distance) -> "Ray":
        raw = type(self).propagate
        if _program.is_device_ray(ray):
            return _program.plan_cache.launch(("propagate", self, distance, ray),
                                              lambda: _program.plan_fn(lambda r: raw(r, distance), ray))
        return raw(ray, distance)

    @staticmethod
    def propagate(ray, distance):
        raise NotImplementedError

    def with_distance(self, distance) -> "Propagator":
        return Propagator(distance, self)


class Propagator(NamedTuple):
    """(distance, propagator) pair that acts on a ray (reference propagator.py:22-40)."""

    distance: float
    propagator: BasePropagator

    def __call__(self, ray: "Ray") -> "Ray":
        return self.propagator(ray, self.distance)


class FreeSpaceParaxial(BasePropagator):
    """Paraxial free-space propagation with constant slopes."""

    @staticmethod
    def propagate(ray, distance):
        if _program.is_device_ray(ray):
            return _program.plan_cache.launch(
                ("paraxial", distance, ray), lambda: _program.plan_fn(lambda r: _propagate_paraxial(r, distance), ray))
        return _propagate_paraxial(ray, distance)


def _sqrt(v):
    """Square root of a scalar, numpy or host-tensor value; per-ray values of a
    traced ray program have none."""
    if isinstance(v, numbers.Real):
        return math.sqrt(v)
    if isinstance(v, (np.ndarray, np.generic)):
        return np.sqrt(v)
    if isinstance(v, torch.Tensor):
        return torch.sqrt(v)
    raise NotImplementedError(
        "FreeSpaceDirCosine needs a square root per ray, which ray programs do not have: propagate scalar or "
        "host rays with it, or use FreeSpaceParaxial on device rays")


class FreeSpaceDirCosine(BasePropagator):
    """Free-space propagation with direction cosines (reference
    propagator.py:79-116): N = sqrt(1 + dx^2 + dy^2), L = dx / N, M = dy / N,
    x += (L / N) d, y += (M / N) d, z = z * _one + d, pathlength += d N; the
    slopes are unchanged.  For scalar and host rays (the square root has no
    ray-program instruction); like the reference's it is not part of the
    ABCD matrices."""

    @staticmethod
    def propagate(ray, distance):
        n = _sqrt(1 + ray.dx**2 + ray.dy**2)
        el = ray.dx / n
        em = ray.dy / n
        return ray.derive(
            x=ray.x + el / n * distance,
            y=ray.y + em / n * distance,
            z=ray.z * ray._one + distance,
            pathlength=ray.pathlength + distance * n,
        )
