"""The composed affine-plus-pathlength form (compose.py) on the host: the
README golden, the algebraic identity with the oracle's stepwise float64
evaluation (to 1e-12 relative, the bar for the composed form), the size of
its ray program, and rejection of non-affine elements."""
import numpy as np
import pytest
import torch

from oracle import ray_oracle as O
from temgymcore_amd import _program
from temgymcore_amd.components import (Biprism, Deflector, DescanError, Descanner, Detector, Lens, Plane, Rotator,
                                       Scanner)
from temgymcore_amd.compose import affine_form, apply_affine
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import run_to_end
from test_host_compiler import interpret


def test_readme_golden_and_matrix():
    """README.md:35-52 and :230-234: Lens(0.5, f=1) -> Detector(1.0)."""
    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))]
    f = affine_form(model)
    np.testing.assert_allclose(f.matrix, [[0.5, 0, 0.75, 0, 0], [0, 0.5, 0, 0.75, 0], [-1, 0, 0.5, 0, 0],
                                          [0, -1, 0, 0.5, 0], [0, 0, 0, 0, 1]], atol=1e-15)
    out = run_to_end(Ray(0.1, 0.2, 0.3, 0.4, 0.0, 0.0), model, composed=True)
    assert out.x == pytest.approx(0.275, abs=1e-15) and out.dx == pytest.approx(0.05, abs=1e-15)
    assert out.pathlength == pytest.approx(0.88875, abs=1e-15) and out.z == 1.0


def _affine_model(err, n_rep=1):
    model, omodel = [], []
    z = 0.0
    for k in range(n_rep):
        z0 = z
        model += [Scanner(z=z0 + 0.1, scan_pos_x=0.003, scan_pos_y=-0.002, scan_tilt_x=1e-3, scan_tilt_y=-2e-3),
                  Deflector(z=z0 + 0.2, def_x=0.01, def_y=-0.02), Lens(z=z0 + 0.35, focal_length=0.7 + 0.01 * k),
                  Rotator(z=z0 + 0.4, angle=12.0 + k),
                  Descanner(z=z0 + 0.5, scan_pos_x=0.003, scan_pos_y=-0.002, scan_tilt_x=1e-3, scan_tilt_y=-2e-3,
                            descan_error=DescanError(*err)), Plane(z=z0 + 0.6)]
        omodel += [("scanner", z0 + 0.1, 0.003, -0.002, 1e-3, -2e-3), ("deflector", z0 + 0.2, 0.01, -0.02),
                   ("lens", z0 + 0.35, 0.7 + 0.01 * k), ("rotator", z0 + 0.4, 12.0 + k),
                   ("descanner", z0 + 0.5, 0.003, -0.002, 1e-3, -2e-3, err), ("plane", z0 + 0.6)]
        z = z0 + 0.6
    model.append(Detector(z=z + 0.4, pixel_size=(0.01, 0.01), shape=(64, 64)))
    omodel.append(("detector", z + 0.4))
    return model, omodel


@pytest.mark.parametrize("n_rep", [1, 4])
def test_identity_with_stepwise_oracle(n_rep):
    """Every element kind the composed form covers, once and repeated (a
    25-element model); outputs within 1e-12 of the size of their terms."""
    rng = np.random.default_rng(71)
    err = tuple(float(v) for v in rng.normal(0, 1e-3, 12))
    model, omodel = _affine_model(err, n_rep)
    n = 2000
    x, y = rng.uniform(-0.05, 0.05, (2, n))
    dx, dy = rng.uniform(-0.02, 0.02, (2, n))
    pl = rng.uniform(0, 1e-3, n)
    form = affine_form(model, z_in=0.05)
    got = apply_affine(form, Ray(x=x, y=y, dx=dx, dy=dy, z=0.05, pathlength=pl))
    ref = O.run_to_end(O.make_ray(x, y, dx, dy, 0.05, pl), omodel)
    scale = 1.0 + np.abs(x) + np.abs(y) + np.abs(dx) + np.abs(dy)
    for k in ("x", "y", "dx", "dy", "pathlength"):
        assert np.all(np.abs(getattr(got, k) - ref[k]) <= 1e-12 * scale * (1 + abs(ref["pathlength"]))), k
    assert got.z == pytest.approx(ref["z"], abs=1e-15)


def test_composed_program_is_short_and_matches_stepwise_program():
    """On traced rays the composed form is one program whatever the model's
    length, and (f64) agrees with the stepwise programs to 1e-12."""
    rng = np.random.default_rng(72)
    err = tuple(float(v) for v in rng.normal(0, 1e-3, 12))
    model, omodel = _affine_model(err, 4)
    n = 1000
    host = [rng.uniform(-0.05, 0.05, n), rng.uniform(-0.05, 0.05, n), rng.uniform(-0.02, 0.02, n),
            rng.uniform(-0.02, 0.02, n)]
    ray = Ray(*(torch.from_numpy(a) for a in host), z=0.0, pathlength=0.0)
    tr = _program.Trace(torch.device("cpu"))
    out = apply_affine(affine_form(model), tr.symbolize(ray))
    vids = [getattr(out, k).vid for k in ("x", "y", "dx", "dy", "pathlength")]
    comp = _program.compile_group(tr, vids, [], (n,))  # one program, no split
    assert len(comp.instrs) <= 70
    vals = interpret(tr, comp)
    ref = O.run_to_end(O.make_ray(*host, 0.0, 0.0), omodel)
    for k, v in zip(("x", "y", "dx", "dy", "pathlength"), vids):
        assert np.abs(vals[v] - ref[k]).max() <= 1e-12 * (1 + np.abs(ref[k]).max()), k


def test_rejects_non_affine_and_per_ray_parameters():
    with pytest.raises(NotImplementedError):
        affine_form([Biprism(z=0.1, def_x=1e-3), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))])
    sp = torch.zeros((4, 1), dtype=torch.float64)
    with pytest.raises(TypeError):
        affine_form([Scanner(z=0.1, scan_pos_x=sp, scan_pos_y=sp), Plane(z=0.2)])
    with pytest.raises(ValueError):
        z = torch.zeros(3, dtype=torch.float64)
        run_to_end(Ray(x=z, y=z, dx=z, dy=z, z=z, pathlength=0.0), [Plane(z=0.2)], composed=True)
