"""Free-space propagators (reference propagator.py:8-72).

``FreeSpaceParaxial`` is the propagator of the batched path: x += dx*d,
y += dy*d, z += d, pathlength += d.  On a vectorised ray a call compiles to a
one-step ray program; inside ``run_to_end`` it is fused with the components.
"""
from __future__ import annotations

from typing import NamedTuple, TYPE_CHECKING

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
