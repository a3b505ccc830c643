"""Small public-API behaviours that the reference's own tests assert, on the
host: the coordinate tuples' to_pixels / to_coords (reference
tests/test_coordinates.py:7-42), Ray.to_vector (tests/test_rays.py:106-115),
no dependency of a propagated z on _one (tests/test_rays.py:42-57), the
scan grid's centring (tests/test_component.py:38-71), a descanner's offset
being the same for every ray (:262-311) and a singular component's
Jacobian (:21-36, :381-395)."""
import dataclasses

import numpy as np
import pytest
import torch

from temgymcore_amd import CoordsXY, CoordXY, PixelsYX, PixelYX
from temgymcore_amd.components import Component, DescanError, Descanner, ScanGrid
from temgymcore_amd.propagator import FreeSpaceParaxial
from temgymcore_amd.ray import Ray
from temgymcore_amd.transfer import custom_jacobian_matrix, jacobian


def test_pixel_and_coordinate_tuples():
    for y, x, kind in ((23, 17, "i"), (23.0, 17.0, "f")):
        pxs = PixelYX(y=y, x=x).to_pixels()
        assert isinstance(pxs, PixelsYX)
        assert pxs.x.shape == (1,) and pxs.y.shape == (1,)
        assert (pxs.x[0], pxs.y[0]) == (x, y)
        assert pxs.x.dtype.kind == kind and pxs.y.dtype.kind == kind
    coords = CoordXY(x=17.0, y=23.0).to_coords()
    assert isinstance(coords, CoordsXY)
    assert coords.x.shape == (1,) and (coords.x[0], coords.y[0]) == (17.0, 23.0)
    assert coords.x.dtype.kind == "f" and coords.y.dtype.kind == "f"


def test_to_vector_gives_arrays():
    v = Ray(x=1.0, y=2.0, dx=0.1, dy=0.2, z=3.0, pathlength=4.0).to_vector()
    for f in dataclasses.fields(v):
        t = getattr(v, f.name)
        assert isinstance(t, torch.Tensor) and t.shape == (1,), f.name
    assert float(v.z[0]) == 3.0 and float(v._one[0]) == 1.0


def test_propagated_z_does_not_depend_on_one():
    """d(z)/d(_one) is exactly zero through a paraxial propagation, while
    x depends on it only as the homogeneous coordinate (not at all here)."""
    jac = jacobian(lambda r: FreeSpaceParaxial().propagate(r, 0.1))(
        Ray(x=0.5, y=-0.5, dx=0.1, dy=-0.2, z=0.0, pathlength=0.0))
    assert float(jac.z._one) == 0.0
    assert float(jac.x.dx) == pytest.approx(0.1, abs=1e-15)


@pytest.mark.parametrize("scan_shape", [(5, 5), (3, 7), (4, 4), (5, 8)])
def test_scan_grid_is_centred(scan_shape):
    """Pixel centres are symmetric about 0: an odd side has 0 at its middle
    pixel, an even side has no 0 and mirror values about its middle."""
    h, w = scan_shape
    grid = ScanGrid(z=0.0, rotation=0.0, pixel_size=(0.1, 0.1), shape=scan_shape)
    ys, xs = np.arange(h), np.arange(w)
    yvals = np.asarray(grid.pixels_to_metres((ys, np.zeros_like(ys))).y, dtype=np.float64)
    xvals = np.asarray(grid.pixels_to_metres((np.zeros_like(xs), xs)).x, dtype=np.float64)
    for n, vals in ((h, yvals), (w, xvals)):
        np.testing.assert_allclose(vals, -vals[::-1], atol=1e-15)
        if n % 2:
            assert vals[n // 2] == pytest.approx(0.0, abs=1e-15)
            assert np.count_nonzero(np.abs(vals) > 1e-15) == n - 1
        else:
            assert np.all(np.abs(vals) > 1e-3)
        np.testing.assert_allclose(np.abs(np.diff(vals)), 0.1, rtol=1e-12)


def test_descanner_offset_is_the_same_for_every_ray():
    rng = np.random.default_rng(11)
    err = DescanError(*rng.uniform(0.0, 1.0, 12))
    px, py = rng.uniform(-5.0, 5.0, 2)
    desc = Descanner(z=11.0, scan_pos_x=float(px), scan_pos_y=float(py), descan_error=err)
    offsets = []
    for x, y, dx, dy in rng.normal(size=(10, 4)):
        r = Ray(x=float(x), y=float(y), dx=float(dx), dy=float(dy), z=0.0, pathlength=0.0)
        out = desc(r)
        offsets.append([out.x - r.x, out.y - r.y, out.dx - r.dx, out.dy - r.dy])
    offsets = np.array(offsets)
    np.testing.assert_allclose(offsets, np.broadcast_to(offsets[0], offsets.shape), atol=1e-12)
    assert np.abs(offsets[0]).max() > 0.0


@dataclasses.dataclass(frozen=True)
class _XIntoY(Component):
    """Copies x into y: the Jacobian has two equal rows."""

    z: float = 0.0

    def __call__(self, ray):
        return Ray(x=ray.x, y=ray.x, dx=ray.dx, dy=ray.dy, _one=ray._one, pathlength=ray.pathlength, z=ray.z)


def test_singular_component_jacobian():
    J = custom_jacobian_matrix(jacobian(_XIntoY())(Ray(x=0.0, y=0.0, dx=1.0, dy=1.0, z=0.0, pathlength=0.0)))
    np.testing.assert_array_equal(J[0], J[1])
    assert np.linalg.matrix_rank(J) == 4
    with pytest.raises(np.linalg.LinAlgError):
        np.linalg.inv(J)
