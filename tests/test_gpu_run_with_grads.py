"""run_with_grads on vectorised rays: one fused launch per variable computes
the outputs and their derivatives; checked against closed forms and the
oracle's f64 central differences (f32 rays: relative 2e-5 of the field's
largest derivative; f64 rays: 1e-8, the roundoff of those differences
on a pathlength of about 1 m)."""
import numpy as np
import pytest
import torch

from oracle import ray_oracle as O
from temgymcore_amd.components import DescanError, Descanner, Deflector, Detector, Lens, Rotator, Scanner
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import run_to_end, run_with_grads

pytestmark = pytest.mark.gpu
FIELDS = ("x", "y", "dx", "dy", "pathlength")


def _rays(device, n=50_000, dtype=np.float32, seed=0):
    x, y, dx, dy = O.plain_workload(n, seed=seed, dtype=dtype)
    t = [torch.from_numpy(a).to(device) for a in (x, y, dx, dy)]
    return Ray(x=t[0], y=t[1], dx=t[2], dy=t[3], z=0.0, pathlength=0.0), (x, y, dx, dy)


def _fd(build, p0, host, h):
    def run(p):
        r = O.make_ray(*(a.astype(np.float64) for a in host), 0.0, 0.0)
        return O.run_to_end(r, build(p))

    def cd(hh):
        a, b = run(p0 + hh), run(p0 - hh)
        return {k: (np.asarray(a[k]) - np.asarray(b[k])) / (2 * hh) for k in FIELDS}
    d1, d2 = cd(h), cd(h / 2)
    return {k: (4 * d2[k] - d1[k]) / 3 for k in FIELDS}


def _close(got, ref, rel, what=""):
    got = got.detach().cpu().numpy().astype(np.float64) if isinstance(got, torch.Tensor) else np.float64(got)
    ref = np.broadcast_to(np.asarray(ref, dtype=np.float64), np.shape(got) or ())
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got - ref).max()
    assert err <= rel * scale + 1e-30, f"{what}: max err {err:.3e}, scale {scale:.3e}, rel {err / scale:.2e}"


def test_lens_focal_length_and_value(device):
    ray, (x, y, dx, dy) = _rays(device)
    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]
    value, grads = run_with_grads(ray, model, [model[0].params.focal_length])
    ref = run_to_end(ray, model)
    for k in FIELDS:
        assert torch.equal(getattr(value, k), getattr(ref, k)), k
    g = grads[(model[0], "focal_length")]
    x1 = x.astype(np.float64) + 0.5 * dx.astype(np.float64)
    y1 = y.astype(np.float64) + 0.5 * dy.astype(np.float64)
    _close(g.x, 0.5 * x1, 2e-6)
    _close(g.dx, x1, 2e-6)
    _close(g.pathlength, 0.5 * (x1 * x1 + y1 * y1), 2e-6)
    assert g.x.dtype == torch.float32 and g.x.shape == ray.x.shape


@pytest.mark.parametrize("dtype, rel", [(np.float32, 2e-5), (np.float64, 1e-8)])
def test_chain_of_elements_vs_oracle(device, dtype, rel):
    ray, host = _rays(device, n=20_000, dtype=dtype, seed=3)
    de = DescanError(pxo_pxi=0.07, pyo_pyi=-0.03, sxo_pxi=0.01, offpxi=1e-4)
    model = [Scanner(z=0.1, scan_pos_x=4e-3, scan_pos_y=1e-3, scan_tilt_x=2e-4),
             Deflector(z=0.25, def_x=2e-3, def_y=-1e-3),
             Rotator(z=0.3, angle=17.0),
             Descanner(z=0.4, scan_pos_x=3e-3, scan_pos_y=-2e-3, descan_error=de),
             Lens(z=0.6, focal_length=0.8),
             Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))]

    def build(**over):
        v = dict(spx=4e-3, defx=2e-3, angle=17.0, pxo=0.07, f=0.8, zl=0.6)
        v.update(over)
        return [("scanner", 0.1, v["spx"], 1e-3, 2e-4, 0.0), ("deflector", 0.25, v["defx"], -1e-3),
                ("rotator", 0.3, v["angle"]),
                ("descanner", 0.4, 3e-3, -2e-3, 0.0, 0.0, tuple(de._replace(pxo_pxi=v["pxo"]))),
                ("lens", v["zl"], v["f"]), ("detector", 1.0)]

    cases = [(model[0].params.scan_pos_x, "spx", 4e-3), (model[1].params.def_x, "defx", 2e-3),
             (model[2].params.angle, "angle", 17.0), (model[3].params.descan_error.pxo_pxi, "pxo", 0.07),
             (model[4].params.focal_length, "f", 0.8), (model[4].params.z, "zl", 0.6)]
    _, grads = run_with_grads(ray, model, [c[0] for c in cases])
    assert len(grads) == len(cases)
    for (var, name, p0), g in zip(cases, grads.values()):
        fd = _fd(lambda p: build(**{name: p}), p0, host, 1e-3 * abs(p0))
        for k in FIELDS:
            _close(getattr(g, k), fd[k], rel, f"d {k} / d {name}")


def test_input_ray_fields_per_ray(device):
    ray, _ = _rays(device, n=1000)
    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))]
    _, grads = run_with_grads(ray, model, [ray.params.x, ray.params.dx])
    gx, gdx = grads[(ray, "x")], grads[(ray, "dx")]
    assert torch.all(gx.x == 0.5) and torch.all(gx.dx == -1.0)
    assert torch.all(gdx.x == 0.75) and torch.all(gdx.dx == 0.5)
    assert gx.x.shape == ray.x.shape
