"""temgymcore_amd: MI355X-native batched propagation for TemGymCore's API.

Drop-in surface (same names and semantics as temgym_core):

* ``Ray``, ``Component`` and the optical elements in ``components``;
* ``FreeSpaceParaxial`` / ``Propagator``;
* ``run_iter``, ``run_to_end`` (+ the fused ``run_to_image``);
* ``Grid`` methods (``metres_to_pixels``, ``pixels_to_metres``, ``coords``,
  ``into_image``) on ``Detector`` / ``ScanGrid``;
* ``gaussian.GaussianRay``, ``make_gaussian_image``,
  ``evaluate_gaussian_input_image``.

Vectorised work runs on hand-written gfx950 HIP kernels through the C ABI of
``libtemgym_hip.so``; there is no CPU execution path for tensor rays.
"""
from ._types import (  # noqa: F401
    CoordsXY, CoordXY, Degrees, NonNegativeFloat, PixelsYX, PixelYX, PositiveFloat, Radians,
    ScaleYX, ShapeYX,
)

__version__ = "0.1.0"
