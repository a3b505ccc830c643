"""AberratedLensKrivanek on the GPU: the fused ray program, beamlet images
through the lens (each beamlet's own 5x5 from forward-mode tangents) and
run_with_grads with respect to an aberration coefficient, against the
oracle's restatement of the reference (aberrations.py, components.py:177-215).

Tolerances: ray values within 1e-12 (f64) or 2e-6 (f32) of the size of the
terms they sum (the lens evaluates the Krivanek series as a polynomial, the
reference through hypot/atan2/cos); beamlet fields as in test_gpu_beamlet
(2e-5 of max|field|, 1e-4 rad)."""
import dataclasses
import json
import os

import numpy as np
import pytest
import torch

from oracle import beamlet_oracle as BO
from oracle import ray_oracle as O
from temgymcore_amd.aberrations import KrivanekCoeffs
from temgymcore_amd.components import AberratedLensKrivanek, Detector, Lens
from temgymcore_amd.gaussian import GaussianRay, make_gaussian_image
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import run_to_end, run_with_grads
from temgymcore_amd.transfer import ray_matrix

pytestmark = pytest.mark.gpu


def _coeffs(rng, scale):
    d = {}
    for f in dataclasses.fields(KrivanekCoeffs):
        d[f.name] = rng.uniform(-np.pi, np.pi) if f.name.startswith("phi") else scale * rng.uniform(-1, 1)
    return d


def _grad_scale(ax, ay, c):
    a = np.hypot(ax, ay)
    orders = {"C10": 1, "C12": 1, "C21": 2, "C23": 2, "C30": 3, "C32": 3, "C34": 3, "C41": 4, "C43": 4,
              "C45": 4, "C50": 5, "C52": 5, "C54": 5, "C56": 5}
    return sum(abs(c.get(k, 0.0)) * (n + 1) * a ** n for k, n in orders.items())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_ray_program_vs_oracle(device, dtype):
    rng = np.random.default_rng(61)
    n, f = 200_000, 2.5e-4
    c = _coeffs(rng, 1e4)
    host = [rng.uniform(-2e-8, 2e-8, n).astype(dtype) for _ in range(2)]
    host += [rng.uniform(-1e-5, 1e-5, n).astype(dtype) for _ in range(2)]
    t = [torch.from_numpy(a).to(device) for a in host]
    model = [AberratedLensKrivanek(z=f, focal_length=f, coeffs=c), Detector(z=2 * f, pixel_size=(1e-6, 1e-6),
                                                                             shape=(8, 8))]
    out = run_to_end(Ray(x=t[0], y=t[1], dx=t[2], dy=t[3], z=0.0, pathlength=0.0), model)
    x, y, dx, dy = (a.astype(np.float64) for a in host)
    ref = O.run_to_end(O.make_ray(x, y, dx, dy), [("kriv", f, f, c), ("detector", 2 * f)])
    tol = 1e-12 if dtype == np.float64 else 2e-6
    idx, idy = -x / f + dx, -y / f + dy
    sd = np.abs(idx) + np.abs(idy) + _grad_scale(idx, idy, c) / f
    for k, scale in (("dx", sd), ("dy", sd), ("x", np.abs(x) + f * sd), ("y", np.abs(y) + f * sd)):
        got = getattr(out, k).cpu().numpy()
        assert got.dtype == dtype
        assert np.all(np.abs(got - ref[k]) <= tol * scale), k


def test_notebook_golden_on_device(device):
    """examples/aberrated_probe.ipynb cell 9: 1e5 fibonacci-spiral rays
    through an f = 0.25 mm lens with C30 = 1e8 m (reference printout)."""
    with open(os.path.join(os.path.dirname(__file__), "golden", "aberrated_probe_rays.json")) as fh:
        g = json.load(fh)
    st = g["setup"]
    x, y = O.fibonacci_spiral(st["n_beams"], st["radius"], alpha=st["alpha"])
    n, F1 = len(x), st["F1"]
    z = torch.zeros(n, dtype=torch.float64, device=device)
    ray = Ray(x=torch.from_numpy(x).to(device), y=torch.from_numpy(y).to(device), dx=z, dy=z, z=z, pathlength=z)
    model = [AberratedLensKrivanek(z=F1, focal_length=F1, coeffs=KrivanekCoeffs(**st["coeffs"])),
             Detector(z=2 * F1, pixel_size=(1e-6, 1e-6), shape=(8, 8))]
    out = run_to_end(ray, model)
    idx = np.array(g["index"])
    for k in ("x", "y", "dx", "dy"):
        np.testing.assert_allclose(getattr(out, k).cpu().numpy()[idx], g[k], rtol=6e-9, atol=0, err_msg=k)


def _abcd_richardson(model_t, x, y, dx, dy):
    """Per-ray 5x5 Jacobians of the oracle by Richardson central differences."""
    n = len(x)
    base = np.stack([x, y, dx, dy, np.ones(n)], 1)
    steps = np.array([1e-10, 1e-10, 1e-9, 1e-9, 1e-6])
    J = np.zeros((n, 5, 5))

    def run(v):
        r = O.run_to_end(O.make_ray(v[:, 0], v[:, 1], v[:, 2], v[:, 3], 0.0, 0.0, v[:, 4]), model_t)
        return np.stack([r["x"], r["y"], r["dx"], r["dy"], np.broadcast_to(r["one"], (n,))], 1)

    for j in range(5):
        def cd(h):
            a, b = base.copy(), base.copy()
            a[:, j] += h
            b[:, j] -= h
            return (run(a) - run(b)) / (2 * h)

        J[:, :, j] = (4 * cd(steps[j] / 2) - cd(steps[j])) / 3
    return J


def _gray(x, y, dx, dy, rng, lam=2e-12, w=5e-8):
    n = len(x)
    amp = rng.uniform(0.5, 1.5, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    d = dict(x=x, y=y, dx=dx, dy=dy, pathlength=rng.uniform(0, 1e-9, n), amplitude=amp,
             waist_xy=w * rng.uniform(0.7, 1.3, (n, 2)), radii=np.full((n, 2), np.inf), wavelength=np.full(n, lam),
             theta=rng.uniform(-1, 1, n))
    t = torch.from_numpy
    g = GaussianRay(x=t(x), y=t(y), dx=t(dx), dy=t(dy), z=torch.zeros(n, dtype=torch.float64),
                    pathlength=t(d["pathlength"]), _one=torch.ones(n, dtype=torch.float64), amplitude=t(amp),
                    waist_xy=t(d["waist_xy"]), radii_of_curv=t(d["radii"]), wavelength=t(d["wavelength"]),
                    theta=t(d["theta"]))
    return g, d


def test_beamlet_image_through_aberrated_lens_vs_oracle(device):
    """Spherical aberration and astigmatism make every beamlet's matrix
    different; the image matches the oracle's sum with per-beamlet
    Jacobians of the oracle's own lens."""
    rng = np.random.default_rng(62)
    n, f = 24, 5e-4
    x, y = rng.uniform(-2e-7, 2e-7, (2, n))
    dx, dy = rng.uniform(-1e-6, 1e-6, (2, n))
    # ~10% changes of the paraxial matrix across the beamlets (slopes ~4e-4)
    c = {"C30": 50.0, "C12": 5e-6, "phi12": 0.6, "C21": 0.03, "phi21": -0.4}
    model = [AberratedLensKrivanek(z=f, focal_length=f, coeffs=c),
             Detector(z=1.5 * f, pixel_size=(3e-8, 3e-8), shape=(72, 80))]
    g, d = _gray(x, y, dx, dy, rng)
    got = make_gaussian_image(g, model).cpu().numpy()
    J = _abcd_richardson([("kriv", f, f, c), ("detector", 1.5 * f)], x, y, dx, dy)
    assert np.abs(J[:, 2, 0] - J[0, 2, 0]).max() > 1e-3 * abs(J[0, 2, 0])  # ray-dependent matrices
    r2 = BO.grid_coords((3e-8, 3e-8), (72, 80))
    ref = BO.make_image(x, y, dx, dy, d["pathlength"], d["amplitude"], d["waist_xy"], d["radii"], d["wavelength"],
                        d["theta"], J, r2).reshape(72, 80)
    got = got.astype(np.complex128)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 2e-5 * scale
    strong = np.abs(ref) > 1e-3 * scale
    assert np.abs(np.angle(got[strong] / ref[strong])).max() <= 1e-4


def test_traced_abcd_matches_host_matrices(device):
    from temgymcore_amd.gaussian import _abcd_traced

    rng = np.random.default_rng(63)
    n = 8
    x, y = rng.uniform(-2e-8, 2e-8, (2, n))
    dx, dy = rng.uniform(-1e-5, 1e-5, (2, n))
    c = _coeffs(rng, 1e4)
    model = [Lens(z=1e-4, focal_length=3e-4), AberratedLensKrivanek(z=2.5e-4, focal_length=2.5e-4, coeffs=c),
             Detector(z=5e-4, pixel_size=(3e-8, 3e-8), shape=(8, 8))]
    g, _ = _gray(x, y, dx, dy, rng)
    J = _abcd_traced(g, model, device, n).cpu().numpy()
    for b in range(n):
        H = np.asarray(ray_matrix(model, Ray(float(x[b]), float(y[b]), float(dx[b]), float(dy[b]), 0.0, 0.0)))
        scale = np.abs(H).max(axis=1, keepdims=True)
        assert np.all(np.abs(J[b] - H) <= 1e-12 * scale + 1e-300)


def test_run_with_grads_aberration_coefficient(device):
    """d(output)/dC30 of device rays against Richardson differences of the oracle."""
    rng = np.random.default_rng(64)
    n, f = 20_000, 2.5e-4
    host = [rng.uniform(-2e-8, 2e-8, n), rng.uniform(-2e-8, 2e-8, n),
            rng.uniform(-1e-5, 1e-5, n), rng.uniform(-1e-5, 1e-5, n)]
    t = [torch.from_numpy(a).to(device) for a in host]
    c = {"C30": 1e6, "C12": 1e-3, "phi12": 0.2}
    lens = AberratedLensKrivanek(z=f, focal_length=f, coeffs=c)
    model = [lens, Detector(z=2 * f, pixel_size=(1e-6, 1e-6), shape=(8, 8))]
    _, grads = run_with_grads(Ray(x=t[0], y=t[1], dx=t[2], dy=t[3], z=0.0, pathlength=0.0), model,
                              [lens.params.coeffs.C30])
    g = grads[(lens, "coeffs", "C30")]

    def run(v):
        return O.run_to_end(O.make_ray(*host, 0.0, 0.0), [("kriv", f, f, dict(c, C30=v)), ("detector", 2 * f)])

    h = 1e3

    def cd(hh):
        a, b = run(c["C30"] + hh), run(c["C30"] - hh)
        return {k: (a[k] - b[k]) / (2 * hh) for k in ("x", "y", "dx", "dy", "pathlength")}

    d1, d2 = cd(h), cd(h / 2)
    for k in ("x", "y", "dx", "dy", "pathlength"):
        ref = (4 * d2[k] - d1[k]) / 3
        got = getattr(g, k).cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-7 * np.abs(ref).max(), k


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(128, 128), (256, 200)])
def test_aberrated_lens_with_binning(device, dtype, shape):
    """Every Krivanek term at once splits the model into two programs; the
    fused binning of the last, into a small (LDS-histogram size) or large
    detector, counts exactly the pixels of the positions the launch returns."""
    from temgymcore_amd.run import run_to_image

    rng = np.random.default_rng(65)
    n, f = 300_000, 2.5e-4
    # every coefficient non-zero, at Krivanek-like sizes (order n ~ 10^(2n-9) m)
    c = {k: (v if k.startswith("phi") else v * 10.0 ** (2 * int(k[1]) - 9)) for k, v in _coeffs(rng, 1.0).items()}
    host = [rng.uniform(-2e-8, 2e-8, n).astype(dtype) for _ in range(2)]
    host += [rng.uniform(-1e-2, 1e-2, n).astype(dtype) for _ in range(2)]
    t = [torch.from_numpy(a).to(device) for a in host]
    det = Detector(z=2 * f, pixel_size=(5e-8, 5e-8), shape=shape)
    model = [AberratedLensKrivanek(z=f, focal_length=f, coeffs=c), det]
    out, img = run_to_image(Ray(x=t[0], y=t[1], dx=t[2], dy=t[3], z=0.0, pathlength=0.0), model)
    x, y = out.x.cpu().numpy(), out.y.cpu().numpy()
    Tinv = O.metres_to_pixels_matrix((0.0, 0.0), (5e-8, 5e-8), shape)
    py, px = O.metres_to_pixels(x, y, Tinv)
    ref = O.into_image(py, px, shape)
    assert ref.sum() > 0.5 * n  # most rays land
    np.testing.assert_array_equal(img.cpu().numpy().astype(np.int64), ref.astype(np.int64))
