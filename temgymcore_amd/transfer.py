"""5x5 ray-transfer matrices (reference transfer.py:6-183, utils.py:7-43, run.py:26-41).

Jacobians of component bodies are taken by forward-mode dual numbers on the
host: the component's own ``__call__`` body is executed on a ray whose fields
carry 7 tangents (x, y, dx, dy, z, pathlength, _one).  For the linear elements
of the batched path the result is exact (the bodies are affine in the ray).
"""
from __future__ import annotations

import dataclasses

import numpy as np

from ._program import RAY_FIELDS, is_number

_N = len(RAY_FIELDS)


class Dual:
    """value + tangent vector; supports the arithmetic used by component bodies."""

    __slots__ = ("v", "g")

    def __init__(self, v, g):
        self.v = float(v)
        self.g = g  # np.ndarray of shape (_N,)

    @staticmethod
    def lift(o):
        return o if isinstance(o, Dual) else Dual(o, np.zeros(_N))

    def __add__(self, o):
        o = Dual.lift(o)
        return Dual(self.v + o.v, self.g + o.g)

    __radd__ = __add__

    def __sub__(self, o):
        o = Dual.lift(o)
        return Dual(self.v - o.v, self.g - o.g)

    def __rsub__(self, o):
        return Dual.lift(o) - self

    def __mul__(self, o):
        o = Dual.lift(o)
        return Dual(self.v * o.v, self.g * o.v + o.g * self.v)

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = Dual.lift(o)
        return Dual(self.v / o.v, (self.g * o.v - o.g * self.v) / (o.v * o.v))

    def __rtruediv__(self, o):
        return Dual.lift(o) / self

    def __neg__(self):
        return Dual(-self.v, -self.g)

    def __pos__(self):
        return self

    def __pow__(self, p):
        if not is_number(p):
            raise NotImplementedError
        return Dual(self.v ** p, self.g * (p * self.v ** (p - 1)) if p != 0 else np.zeros(_N))

    def sign(self):
        s = 0.0 if self.v == 0 else (1.0 if self.v > 0 else -1.0)
        return Dual(s, np.zeros(_N))

    def __repr__(self):
        return f"Dual({self.v}, {self.g})"


def _seed(ray):
    vals = {}
    for f in dataclasses.fields(ray):
        v = getattr(ray, f.name)
        if f.name in RAY_FIELDS:
            g = np.zeros(_N)
            g[RAY_FIELDS.index(f.name)] = 1.0
            vals[f.name] = Dual(float(v), g)
        else:
            vals[f.name] = v
    from ._program import _rebuild

    return _rebuild(ray, vals)


def jacobian(fn, has_aux_value: bool = False):
    """d fn(ray) / d ray for a scalar ray, as a Ray of Rays.

    ``jac.x.dx`` is d(out.x)/d(in.dx), as with ``jax.jacobian`` over a Ray
    pytree.  With `has_aux_value` the call returns (jac, value).
    """
    from .ray import Ray

    def inner(ray):
        out = fn(_seed(ray))
        rows = {}
        value = {}
        for f in dataclasses.fields(out):
            v = getattr(out, f.name)
            if f.name not in RAY_FIELDS:
                value[f.name] = v
                continue
            d = v if isinstance(v, Dual) else Dual.lift(v)
            value[f.name] = d.v
            rows[f.name] = Ray(**{n: float(d.g[k]) for k, n in enumerate(RAY_FIELDS)})
        from ._program import _rebuild

        val = _rebuild(out, value)
        jac = Ray(**rows)
        return (jac, val) if has_aux_value else jac

    return inner


def custom_jacobian_matrix(ray_jac) -> np.ndarray:
    """Nested Ray-Jacobian -> 5x5 homogeneous matrix over (x, y, dx, dy, _one) (reference utils.py:7-43)."""
    keys = ("x", "y", "dx", "dy", "_one")
    return np.array([[float(getattr(getattr(ray_jac, r), c)) for c in keys] for r in keys])


def ray_matrix(model, ray=None, propagator=None) -> np.ndarray:
    """5x5 ABCD of ``run_to_end`` over `model` at the (scalar) central `ray`."""
    from .propagator import FreeSpaceParaxial
    from .ray import Ray
    from .run import run_to_end

    if ray is None:
        ray = Ray.origin()
    prop = FreeSpaceParaxial() if propagator is None else propagator
    return custom_jacobian_matrix(jacobian(lambda r: run_to_end(r, model, prop))(ray))


def transfer_rays(ray_coords, transfer_matrices):
    """Rays (N,5) through cumulative products of (M,5,5) matrices -> (N,M,5) (reference transfer.py:6-54)."""
    cumulative = accumulate_matrices_cumulative(transfer_matrices)
    return np.einsum("mij,nj->nmi", cumulative, ray_coords, optimize=True)


def transfer_rays_pt_src(input_pos_xy, input_slopes_xy, transfer_matrix):
    """One 5x5 matrix applied to point-source rays -> (4, N) (reference transfer.py:57-123)."""
    x0, y0 = input_pos_xy
    sx, sy = (np.asarray(s, dtype=np.float64) for s in input_slopes_xy)
    rays = np.vstack([np.full(sx.shape[0], x0), np.full(sy.shape[0], y0), sx, sy, np.ones_like(sx)])
    end = np.dot(np.asarray(transfer_matrix, dtype=np.float64), rays)
    return np.array([end[0, :], end[1, :], end[2, :], end[3, :]])


def accumulate_matrices(matrices):
    """matrices[-1] @ ... @ matrices[0] (reference transfer.py:126-147)."""
    total = np.asarray(matrices[-1])
    for tm in reversed(matrices[:-1]):
        total = total @ np.asarray(tm)
    return total


def accumulate_matrices_cumulative(matrices):
    """Cumulative products in the reference's order (reference transfer.py:150-183)."""
    out = []
    total = np.asarray(matrices[-1])
    out.append(total)
    for tm in reversed(list(matrices[:-1])):
        total = np.asarray(tm) @ total
        out.append(total)
    return np.stack(out, axis=0)
