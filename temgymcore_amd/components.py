"""Optical elements of the batched path (reference components.py:12-559).

Each component's ``__call__`` body is the plain per-ray arithmetic of the
reference element.  The same body runs three ways:

* on a scalar ray (Python floats): directly, in double precision;
* on a vectorised ray (tensor fields): traced into a ray program and run by
  the HIP ray-program kernel (``Component.__init_subclass__`` installs the
  dispatch, so user-defined subclasses get it too);
* on dual numbers, to obtain 5x5 ray-transfer matrices (``transfer.jacobian``).
"""
from __future__ import annotations

import dataclasses
import functools
import math
from typing import NamedTuple

from . import _program
from ._types import CoordsXY, ScaleYX, ShapeYX
from .grid import Grid
from .ray import Ray
from .tree_utils import HasParamsMixin


def _install_dispatch(cls) -> None:
    raw = cls.__dict__.get("__call__")
    if raw is None or getattr(raw, "_tg_raw", None) is not None:
        return

    @functools.wraps(raw)
    def __call__(self, ray, *args, **kwargs):
        if _program.is_device_ray(ray):
            return _program.plan_cache.launch(
                ("component", raw, self, args, kwargs, ray),
                lambda: _program.plan_fn(lambda r: raw(self, r, *args, **kwargs), ray))
        return raw(self, ray, *args, **kwargs)

    __call__._tg_raw = raw
    cls.__call__ = __call__


class Component(HasParamsMixin):
    """Base component: pure ``__call__(ray) -> Ray`` at axial position ``z``.

    Subclasses write the per-ray update with ordinary arithmetic on the ray's
    fields; it is compiled for the GPU automatically.  Components do not move
    ``ray.z`` -- propagators do (reference components.py:12-24) -- except
    ThickLens, which hands the ray on from its object to its image plane.
    """

    def __init_subclass__(cls, **kw):
        super().__init_subclass__(**kw)
        _install_dispatch(cls)

    def __call__(self, ray: Ray) -> Ray:
        raise NotImplementedError


class DescanError(NamedTuple):
    """Linear descan error coefficients (reference components.py:27-115).

    pos/slope offsets at the output as a linear function of scan position:
    pos_offset_x = pxo_pxi*spx + pxo_pyi*spy + offpxi, etc.
    """

    pxo_pxi: float = 0.0
    pxo_pyi: float = 0.0
    pyo_pxi: float = 0.0
    pyo_pyi: float = 0.0
    sxo_pxi: float = 0.0
    sxo_pyi: float = 0.0
    syo_pxi: float = 0.0
    syo_pyi: float = 0.0
    offpxi: float = 0.0
    offpyi: float = 0.0
    offsxi: float = 0.0
    offsyi: float = 0.0

    def as_array(self):
        import numpy as np

        return np.array(self, dtype=np.float64)

    def as_matrix(self):
        """5x5 matrix with the offsets in the 5th column.

        The reference (components.py:107-115) places ``offsyi`` in the slope-x
        row; this build uses ``offsxi`` there, matching the Descanner maths.
        """
        import numpy as np

        return np.array([
            [self.pxo_pxi, self.pxo_pyi, 0.0, 0.0, self.offpxi],
            [self.pyo_pxi, self.pyo_pyi, 0.0, 0.0, self.offpyi],
            [self.sxo_pxi, self.sxo_pyi, 0.0, 0.0, self.offsxi],
            [self.syo_pxi, self.syo_pyi, 0.0, 0.0, self.offsyi],
            [0.0, 0.0, 0.0, 0.0, 1.0],
        ])


@dataclasses.dataclass(frozen=True)
class Plane(Component):
    """No-op at plane z (reference components.py:118-134)."""

    z: float

    def __call__(self, ray: Ray):
        return ray


@dataclasses.dataclass(frozen=True)
class Lens(Component):
    """Thin lens: dx' = -x/f + dx, pathlength -= (x^2+y^2)/(2f) (reference components.py:137-174)."""

    z: float
    focal_length: float

    def __call__(self, ray: Ray):
        f = self.focal_length
        x, y, dx, dy = ray.x, ray.y, ray.dx, ray.dy
        new_dx = -x / f + dx
        new_dy = -y / f + dy
        pathlength = ray.pathlength - (x**2 + y**2) / (2 * f)
        one = ray._one * 1.0
        return Ray(x=x, y=y, dx=new_dx, dy=new_dy, _one=one, pathlength=pathlength, z=ray.z)


@dataclasses.dataclass(frozen=True)
class ThickLens(Component):
    """Thick lens between an object plane z_po and an image plane z_pi: the
    thin-lens update of the slopes and pathlength at z_po, after which the
    ray continues from z_pi (its z moves by z_pi - z_po).  Its position
    ``z`` is z_po (reference components.py:410-451)."""

    z_po: float
    z_pi: float
    focal_length: float

    @property
    def z(self):
        return self.z_po

    def __call__(self, ray: Ray):
        f = self.focal_length
        x, y, dx, dy = ray.x, ray.y, ray.dx, ray.dy
        new_dx = -x / f + dx
        new_dy = -y / f + dy
        pathlength = ray.pathlength - (x**2 + y**2) / (2 * f)
        new_z = ray.z - (self.z_po - self.z_pi)
        one = ray._one * 1.0
        return Ray(x=x, y=y, dx=new_dx, dy=new_dy, _one=one, pathlength=pathlength, z=new_z)


@dataclasses.dataclass(frozen=True)
class AberratedLensKrivanek(Lens):
    """Thin lens with Krivanek aberrations (reference components.py:177-215).

    The paraxial slopes a = (-x/f + dx, -y/f + dy) are deflected by -grad W(a)/f
    and the pathlength gains W(a)/f, with W the Krivanek series of ``coeffs``
    (a KrivanekCoeffs, or a dict of its fields).  W and its gradient are
    evaluated as polynomials in the slopes (aberrations.krivanek_poly), so
    the per-ray program needs no transcendental functions."""

    coeffs: object = None

    def __post_init__(self):
        from .aberrations import as_coeffs

        object.__setattr__(self, "coeffs", as_coeffs(self.coeffs if self.coeffs is not None else {}))

    def __call__(self, ray: Ray):
        from .aberrations import krivanek_poly

        f = self.focal_length
        x, y, dx, dy = ray.x, ray.y, ray.dx, ray.dy
        ideal_dx = -x / f + dx
        ideal_dy = -y / f + dy
        W, dWx, dWy = krivanek_poly(ideal_dx, ideal_dy, self.coeffs)
        aber_dx = ideal_dx + (-dWx / f)
        aber_dy = ideal_dy + (-dWy / f)
        pathlength = ray.pathlength - (x**2 + y**2) / (2 * f) + W / f
        one = ray._one * 1.0
        return Ray(x=x, y=y, dx=aber_dx, dy=aber_dy, _one=one, pathlength=pathlength, z=ray.z)


@dataclasses.dataclass(frozen=True)
class ScanGrid(Component, Grid):
    """Scan grid: pixel<->metre mapping at plane z; a no-op on rays (reference components.py:218-249)."""

    z: float
    pixel_size: ScaleYX
    shape: ShapeYX
    rotation: float = 0.0
    centre: CoordsXY = (0.0, 0.0)
    flip_y: bool = False

    def __call__(self, ray: Ray):
        return ray


@dataclasses.dataclass(frozen=True)
class Scanner(Component):
    """Adds scan position and tilt offsets (reference components.py:252-285)."""

    z: float
    scan_pos_x: float
    scan_pos_y: float
    scan_tilt_x: float = 0.0
    scan_tilt_y: float = 0.0

    def __call__(self, ray: Ray):
        return ray.derive(
            x=ray.x + self.scan_pos_x * ray._one,
            y=ray.y + self.scan_pos_y * ray._one,
            dx=ray.dx + self.scan_tilt_x * ray._one,
            dy=ray.dy + self.scan_tilt_y * ray._one,
        )


@dataclasses.dataclass(frozen=True)
class Descanner(Component):
    """Undoes the scan with a linear descan error (reference components.py:288-372)."""

    z: float
    scan_pos_x: float
    scan_pos_y: float
    scan_tilt_x: float = 0.0
    scan_tilt_y: float = 0.0
    descan_error: DescanError = DescanError()

    def __call__(self, ray: Ray):
        de = self.descan_error
        sp_x, sp_y = self.scan_pos_x, self.scan_pos_y
        st_x, st_y = self.scan_tilt_x, self.scan_tilt_y
        return ray.derive(
            x=ray.x + (sp_x * de.pxo_pxi + sp_y * de.pxo_pyi + de.offpxi - sp_x) * ray._one,
            y=ray.y + (sp_x * de.pyo_pxi + sp_y * de.pyo_pyi + de.offpyi - sp_y) * ray._one,
            dx=ray.dx + (sp_x * de.sxo_pxi + sp_y * de.sxo_pyi + de.offsxi - st_x) * ray._one,
            dy=ray.dy + (sp_x * de.syo_pxi + sp_y * de.syo_pyi + de.offsyi - st_y) * ray._one,
        )


@dataclasses.dataclass(frozen=True)
class Detector(Component, Grid):
    """Detector grid at plane z; a no-op on rays (reference components.py:375-406)."""

    z: float
    pixel_size: ScaleYX
    shape: ShapeYX
    rotation: float = 0.0
    centre: CoordsXY = (0.0, 0.0)
    flip_y: bool = False

    def __call__(self, ray: Ray):
        return ray


@dataclasses.dataclass(frozen=True)
class Deflector(Component):
    """Constant deflection; pathlength += dx_in*x + dy_in*y (reference components.py:455-482).

    The pathlength term uses the incoming slopes, exactly as the reference.
    """

    z: float
    def_x: float
    def_y: float

    def __call__(self, ray: Ray):
        x, y, dx, dy = ray.x, ray.y, ray.dx, ray.dy
        return ray.derive(
            dx=dx + self.def_x * ray._one,
            dy=dy + self.def_y * ray._one,
            pathlength=ray.pathlength + dx * x + dy * y,
        )


@dataclasses.dataclass(frozen=True)
class Rotator(Component):
    """Rotates positions and slopes about the axis by `angle` degrees (reference components.py:485-523)."""

    z: float
    angle: float

    def __call__(self, ray: Ray):
        # jnp.deg2rad(a) = a * (pi / 180); cos/sin of a scalar are host f64
        angle = self.angle * (math.pi / 180)
        c, s = _program.cos(angle), _program.sin(angle)
        new_x = ray.x * c - ray.y * s
        new_y = ray.x * s + ray.y * c
        new_dx = ray.dx * c - ray.dy * s
        new_dy = ray.dx * s + ray.dy * c
        return Ray(x=new_x, y=new_y, dx=new_dx, dy=new_dy, _one=ray._one, pathlength=ray.pathlength, z=ray.z)


@dataclasses.dataclass(frozen=True)
class Biprism(Component):
    """dx += def_x * sign(x); pathlength += dx*x + dy*y (reference components.py:526-559).

    As in the reference, `offset` and `rotation` are carried but not applied.
    """

    z: float
    offset: float = 0.0
    rotation: float = 0.0
    def_x: float = 0.0
    side: int = 1

    def __call__(self, ray: Ray):
        x, y, dx, dy = ray.x, ray.y, ray.dx, ray.dy
        return ray.derive(
            dx=dx + self.def_x * ray._one * _program.sign(ray.x),
            dy=dy,
            pathlength=ray.pathlength + dx * x + dy * y,
        )
