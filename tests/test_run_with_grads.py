"""run_with_grads / PathBuilder (reference run.py:182-267, tree_utils.py) on
scalar rays: closed forms of the lens model and central differences of the
oracle's f64 model for every other element."""
import math

import numpy as np
import pytest

from oracle import ray_oracle as O
from temgymcore_amd.components import (DescanError, Descanner, Deflector, Detector, Lens, Rotator, Scanner)
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import run_to_end, run_with_grads
from temgymcore_amd.tree_utils import PathBuilder

RAY = Ray(x=0.01, y=-0.02, dx=0.003, dy=0.001, z=0.0, pathlength=0.0)
FIELDS = ("x", "y", "dx", "dy", "z", "pathlength")


def _lens_model(f=1.0, zl=0.5):
    return [Lens(z=zl, focal_length=f), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]


def test_lens_focal_length_closed_form():
    model = _lens_model()
    value, grads = run_with_grads(RAY, model, [model[0].params.focal_length])
    assert list(grads) == [(model[0], "focal_length")]
    g = grads[(model[0], "focal_length")]
    f, x, y, dx, dy = 1.0, RAY.x, RAY.y, RAY.dx, RAY.dy
    x1, y1 = x + 0.5 * dx, y + 0.5 * dy
    assert g.x == pytest.approx(0.5 * x1 / f**2, rel=1e-14)
    assert g.dx == pytest.approx(x1 / f**2, rel=1e-14)
    assert g.y == pytest.approx(0.5 * y1 / f**2, rel=1e-14)
    assert g.pathlength == pytest.approx((x1**2 + y1**2) / (2 * f**2), rel=1e-13)
    assert g.z == 0.0
    ref = run_to_end(RAY, model)
    assert value == ref


def test_input_ray_fields_closed_form():
    model = _lens_model()
    _, grads = run_with_grads(RAY, model, [RAY])
    assert set(k[1] for k in grads) == {"x", "y", "dx", "dy", "z", "pathlength", "_one"}
    gx = grads[(RAY, "x")]
    assert gx.x == pytest.approx(0.5)  # d/dx [x1 + 0.5(-x1/f + dx)], x1 = x + 0.5 dx
    assert gx.dx == pytest.approx(-1.0)
    gdx = grads[(RAY, "dx")]
    assert gdx.x == pytest.approx(0.5 + 0.5 * (1 - 0.5))
    assert grads[(RAY, "pathlength")].pathlength == pytest.approx(1.0)


def _oracle_fd(build, p0, h):
    """Richardson-extrapolated central differences of the oracle's f64 model."""
    def run(p):
        r = O.make_ray(RAY.x, RAY.y, RAY.dx, RAY.dy, RAY.z, RAY.pathlength)
        return O.run_to_end(r, build(p))

    def cd(hh):
        a, b = run(p0 + hh), run(p0 - hh)
        return {k: (a[k] - b[k]) / (2 * hh) for k in FIELDS}
    d1, d2 = cd(h), cd(h / 2)
    return {k: (4 * d2[k] - d1[k]) / 3 for k in FIELDS}


@pytest.mark.parametrize("case", ["lens_z", "deflector", "rotator", "scanner", "descan"])
def test_elements_vs_oracle_central_differences(case):
    det = Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))
    if case == "lens_z":
        p0 = 0.37
        model = _lens_model(zl=p0)
        var = model[0].params.z
        build = lambda p: [("lens", p, 1.0), ("detector", 1.0)]  # noqa: E731
    elif case == "deflector":
        p0 = 2e-3
        model = [Deflector(z=0.3, def_x=p0, def_y=-1e-3), det]
        var = model[0].params.def_x
        build = lambda p: [("deflector", 0.3, p, -1e-3), ("detector", 1.0)]  # noqa: E731
    elif case == "rotator":
        p0 = 17.0
        model = [Rotator(z=0.2, angle=p0), det]
        var = model[0].params.angle
        build = lambda p: [("rotator", 0.2, p), ("detector", 1.0)]  # noqa: E731
    elif case == "scanner":
        p0 = 4e-3
        model = [Scanner(z=0.1, scan_pos_x=p0, scan_pos_y=1e-3, scan_tilt_x=2e-4), det]
        var = model[0].params.scan_pos_x
        build = lambda p: [("scanner", 0.1, p, 1e-3, 2e-4, 0.0), ("detector", 1.0)]  # noqa: E731
    else:
        p0 = 0.07
        de = DescanError(pxo_pxi=p0, pyo_pyi=-0.03, sxo_pxi=0.01, offpxi=1e-4)
        model = [Descanner(z=0.2, scan_pos_x=3e-3, scan_pos_y=-2e-3, descan_error=de), det]
        var = model[0].params.descan_error.pxo_pxi
        build = lambda p: [("descanner", 0.2, 3e-3, -2e-3, 0.0, 0.0, tuple(de._replace(pxo_pxi=p))),  # noqa: E731
                           ("detector", 1.0)]
    _, grads = run_with_grads(RAY, model, [var])
    (g,) = grads.values()
    fd = _oracle_fd(build, p0, 1e-3 * max(abs(p0), 1e-3))
    for k in FIELDS:
        assert getattr(g, k) == pytest.approx(fd[k], rel=1e-6, abs=1e-12), k


def test_non_leaf_path_expands_to_every_scalar():
    de = DescanError(pxo_pxi=0.01, offpxi=1e-4)
    model = [Descanner(z=0.2, scan_pos_x=3e-3, scan_pos_y=-2e-3, descan_error=de),
             Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))]
    _, grads = run_with_grads(RAY, model, [model[0].params.descan_error])
    assert len(grads) == len(DescanError._fields)
    assert all(k[:2] == (model[0], "descan_error") for k in grads)
    # the x offset moves out.x one to one
    assert grads[(model[0], "descan_error", "offpxi")].x == pytest.approx(1.0)


def test_missing_variable_raises():
    model = _lens_model()
    stranger = Lens(z=0.1, focal_length=2.0)
    with pytest.raises(RuntimeError):
        run_with_grads(RAY, model, [stranger.params.focal_length])


def test_pathbuilder_paths():
    lens = Lens(z=0.5, focal_length=1.0)
    pb = lens.params.focal_length
    assert isinstance(pb, PathBuilder)
    assert pb._path() == (lens, "focal_length")
    assert pb._value() == 1.0
    det = Detector(z=1.0, pixel_size=(0.01, 0.02), shape=(8, 8))
    assert det.params.pixel_size[1]._value() == 0.02
    assert det.params.pixel_size[1]._path() == (det, "pixel_size", 1)
    assert RAY.params.dx._root() is RAY
