"""Composed affine-plus-pathlength form of an affine model.

A model of affine elements (propagation, Lens, Deflector, Scanner,
Descanner, Rotator, Plane, Detector, ScanGrid, sources) maps the ray state
s = (x, y, dx, dy) to

    s' = A s + b,   pathlength' = pathlength + c + g.s + s^T Q s   (Q upper triangular)

whatever its length.  ``affine_form`` obtains A, b, c, g, Q by running the
element bodies once on the host over polynomials of degree <= 2 in
(x, y, dx, dy, pathlength) with float64 coefficients; the composed form
equals the stepwise model as an algebraic identity (the reference's step
order, run.py:47-116; the composition of transfer.py:126-183 extended by the
pathlength).  ``apply_affine`` evaluates it on a ray: on tensor rays that is
one ray program of at most ~65 instructions for any number of elements,
where the stepwise program grows with the model and splits into several
launches.  It rounds differently from the stepwise evaluation (to about 1e-15
relative in float64, 1e-7 in float32), so it is an opt-in mode
(``run_to_end(..., composed=True)``), not the default.
"""
from __future__ import annotations

import dataclasses
import numbers

import numpy as np

from .propagator import BasePropagator, FreeSpaceParaxial
from .ray import Ray

NVAR = 5  # x, y, dx, dy, pathlength_in
_NAMES = ("x", "y", "dx", "dy", "pathlength")


def _is_scalar(v) -> bool:
    if isinstance(v, bool):
        return False
    if isinstance(v, numbers.Number):
        return True
    if isinstance(v, np.ndarray) and v.size == 1:
        return True
    try:
        import torch

        return isinstance(v, torch.Tensor) and v.numel() == 1
    except ImportError:  # pragma: no cover
        return False


class Poly:
    """c + g.v + v^T Q v over v = (x, y, dx, dy, pathlength_in), float64;
    degree above 2, division by a polynomial and non-scalar operands raise."""

    __slots__ = ("c", "g", "Q")

    def __init__(self, c=0.0, g=None, Q=None):
        self.c = float(c)
        self.g = np.zeros(NVAR) if g is None else g
        self.Q = np.zeros((NVAR, NVAR)) if Q is None else Q

    @staticmethod
    def var(i: int) -> "Poly":
        g = np.zeros(NVAR)
        g[i] = 1.0
        return Poly(0.0, g)

    @staticmethod
    def lift(o) -> "Poly":
        if isinstance(o, Poly):
            return o
        if _is_scalar(o):
            return Poly(float(o))
        raise TypeError(f"composed form: per-ray or array operand {type(o).__name__} (only scalar parameters)")

    @property
    def degree(self) -> int:
        if np.any(self.Q != 0):
            return 2
        return 1 if np.any(self.g != 0) else 0

    def __add__(self, o):
        o = Poly.lift(o)
        return Poly(self.c + o.c, self.g + o.g, self.Q + o.Q)

    __radd__ = __add__

    def __neg__(self):
        return Poly(-self.c, -self.g, -self.Q)

    def __sub__(self, o):
        return self + (-Poly.lift(o))

    def __rsub__(self, o):
        return Poly.lift(o) - self

    def __mul__(self, o):
        o = Poly.lift(o)
        if self.degree + o.degree > 2:
            raise NotImplementedError("composed form: the model is not affine in the ray (degree > 2)")
        c = self.c * o.c
        g = self.c * o.g + o.c * self.g
        F = np.outer(self.g, o.g)  # (g1.v)(g2.v) = sum_ab F_ab v_a v_b, folded onto a <= b
        U = np.triu(F + F.T)
        np.fill_diagonal(U, np.diag(F))
        Q = self.c * o.Q + o.c * self.Q + U
        return Poly(c, g, Q)

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, Poly):
            if o.degree > 0:
                raise NotImplementedError("composed form: division by a per-ray value")
            o = o.c
        if not _is_scalar(o):
            raise TypeError("composed form: division by a non-scalar")
        return self * (1.0 / float(o))

    def __pow__(self, p):
        if p == 2:
            return self * self
        if p == 1:
            return self
        raise NotImplementedError("composed form: powers other than 1 and 2")

    def sign(self):
        raise NotImplementedError("composed form: sign() of a ray value (the element is not affine)")


@dataclasses.dataclass
class AffineForm:
    """s' = A s + b (A [4, 4], b [4]); pathlength' = pathlength + c + g.s + s^T Q s
    (g [4], Q [4, 4] upper triangular); z_out; one' = one."""

    A: np.ndarray
    b: np.ndarray
    c: float
    g: np.ndarray
    Q: np.ndarray
    z_out: float

    @property
    def matrix(self) -> np.ndarray:
        """The 5x5 homogeneous ray-transfer matrix (transfer.ray_matrix layout)."""
        M = np.eye(5)
        M[:4, :4] = self.A
        M[:4, 4] = self.b
        return M


def affine_form(model, z_in: float = 0.0, propagator: BasePropagator = FreeSpaceParaxial()) -> AffineForm:
    """Compose `model` for rays entering at z = z_in.  Raises
    NotImplementedError for non-affine elements (Biprism, aberrated lenses) and
    TypeError for per-ray (tensor) element parameters."""
    from .run import _trace_model

    ray = Ray(x=Poly.var(0), y=Poly.var(1), dx=Poly.var(2), dy=Poly.var(3), z=float(z_in),
              pathlength=Poly.var(4), _one=1.0)
    out = _trace_model(ray, list(model), propagator)
    A = np.zeros((4, 4))
    b = np.zeros(4)
    for i, name in enumerate(("x", "y", "dx", "dy")):
        p = Poly.lift(getattr(out, name))
        if p.degree > 1 or p.g[4] != 0:
            raise NotImplementedError(f"composed form: output {name} is not affine in the ray")
        A[i] = p.g[:4]
        b[i] = p.c
    pl = Poly.lift(out.pathlength)
    if pl.g[4] != 1.0 or np.any(pl.Q[4] != 0) or np.any(pl.Q[:, 4] != 0):
        raise NotImplementedError("composed form: the pathlength is not the input pathlength plus a quadratic")
    one = out._one
    if not (_is_scalar(one) and float(one) == 1.0):
        raise NotImplementedError("composed form: the homogeneous coordinate changes")
    z = out.z
    if not _is_scalar(z):
        raise NotImplementedError("composed form: per-ray z")
    return AffineForm(A=A, b=b, c=pl.c, g=pl.g[:4].copy(), Q=pl.Q[:4, :4].copy(), z_out=float(z))


def apply_affine(form: AffineForm, ray: Ray) -> Ray:
    """Evaluate the composed form on a ray (scalars, arrays or traced per-ray
    values): each output is one sum of products, zero coefficients omitted."""
    s = (ray.x, ray.y, ray.dx, ray.dy)

    def lin(coef, const):
        acc = None
        for j in range(4):
            if coef[j] != 0.0:
                term = s[j] * float(coef[j])
                acc = term if acc is None else acc + term
        if const != 0.0 or acc is None:
            acc = float(const) if acc is None else acc + float(const)
        return acc

    xs = [lin(form.A[i], form.b[i]) for i in range(4)]
    # c + sum_i s_i (g_i + sum_{j >= i} Q_ij s_j)
    q = float(form.c)
    for i in range(4):
        row = np.zeros(4)
        row[i:] = form.Q[i, i:]
        if form.g[i] == 0.0 and not np.any(row != 0.0):
            continue
        q = q + s[i] * lin(row, form.g[i])
    return Ray(x=xs[0], y=xs[1], dx=xs[2], dy=xs[3], z=form.z_out, pathlength=ray.pathlength + q, _one=ray._one)
