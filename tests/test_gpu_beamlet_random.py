"""Seeded random beamlet sets for both tile sums against the complex128 oracle.

Each case draws 1-300 beamlets (waists, spread, tilts, curvature, complex
amplitudes; separable -- theta = 0 -- or rotated astigmatic), a model (free
propagation, a lens, or scanner + deflector offsets) at a random distance, a
detector with random pixel size and a shape that is mostly not a multiple of
the 64-pixel tile (some rotated: a separable beamlet is then not separable in
pixel coordinates), and the tile sum the library picks or the VALU recurrence
forced through the tuning knob.  Bar, as for the fixed cases:
|field - ref| <= 2e-5 max|ref| and phase within 1e-4 rad where
|ref| > 1e-3 max|ref|."""
import numpy as np
import pytest

from oracle import beamlet_oracle as BO
from temgymcore_amd import autotune
from temgymcore_amd.components import Deflector, Detector, Lens, Scanner
from temgymcore_amd.gaussian import GaussianRay, make_gaussian_image
from temgymcore_amd.ray import Ray
from temgymcore_amd.transfer import ray_matrix

import torch

pytestmark = pytest.mark.gpu


def _check(got, ref, amp_tol=2e-5, phase_tol=1e-4):
    got = got.astype(np.complex128)
    scale = np.abs(ref).max()
    err = np.abs(got - ref)
    assert err.max() <= amp_tol * scale, f"max |d| {err.max() / scale:.3e} of max|ref|"
    strong = np.abs(ref) > 1e-3 * scale
    dphi = np.angle(got[strong] / ref[strong])
    assert np.abs(dphi).max() <= phase_tol, f"max phase error {np.abs(dphi).max():.3e} rad"


def _gray(d):
    n = len(d["x"])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    return GaussianRay(x=t(d["x"]), y=t(d["y"]), dx=t(d["dx"]), dy=t(d["dy"]), z=torch.zeros(n, dtype=torch.float64),
                       pathlength=t(d["pathlength"]), _one=torch.ones(n, dtype=torch.float64),
                       amplitude=t(np.asarray(d["amplitude"])), waist_xy=t(d["waist_xy"]),
                       radii_of_curv=t(d["radii"]), wavelength=t(d["wavelength"]), theta=t(d["theta"]))


def _random_case(seed):
    rng = np.random.default_rng(4000 + seed)
    n = int(rng.choice([1, 2, 17, int(rng.integers(3, 300))]))
    lam = float(rng.uniform(1.5e-12, 3e-12))
    w = float(rng.uniform(3e-8, 8e-8))
    spread, tilt = float(rng.uniform(0.0, 3e-7)), float(rng.uniform(0.0, 1.5e-5))
    separable = bool(rng.uniform() < 0.5)
    curved = bool(rng.uniform() < 0.6)
    x, y = rng.uniform(-spread, spread, (2, n))
    dx, dy = rng.uniform(-tilt, tilt, (2, n))
    amp = rng.uniform(0.3, 1.5, n) * np.exp(1j * rng.uniform(-np.pi, np.pi, n))
    radii = rng.uniform(0.3, 3.0, (n, 2)) * rng.choice([-1.0, 1.0], (n, 2)) if curved else np.full((n, 2), np.inf)
    if curved:
        radii[rng.uniform(size=n) < 0.3] = np.inf
    d = dict(x=x, y=y, dx=dx, dy=dy, pathlength=rng.uniform(0.0, 1e-9, n), amplitude=amp,
             waist_xy=w * rng.uniform(0.7, 1.3, (n, 2)), radii=radii, wavelength=np.full(n, lam),
             theta=np.zeros(n) if separable else rng.uniform(-1.5, 1.5, n))
    z_det = float(rng.uniform(5e-4, 2e-3))
    pix = float(rng.uniform(2.5e-8, 5e-8))
    shape = (int(rng.integers(20, 260)), int(rng.integers(20, 260)))
    rotation = float(rng.uniform(-40.0, 40.0)) if rng.uniform() < 0.25 else 0.0
    det = Detector(z=z_det, pixel_size=(pix, pix), shape=shape, rotation=rotation)
    kind = ("free", "lens", "offsets")[int(rng.integers(3))]
    if kind == "free":
        model = [det]
    elif kind == "lens":
        model = [Lens(z=float(rng.uniform(0.2, 0.6)) * z_det, focal_length=float(rng.uniform(0.3, 1.0)) * z_det), det]
    else:
        model = [Scanner(z=0.1 * z_det, scan_pos_x=float(rng.normal(0, 2e-7)), scan_pos_y=float(rng.normal(0, 2e-7))),
                 Deflector(z=0.2 * z_det, def_x=float(rng.normal(0, 1e-5)), def_y=float(rng.normal(0, 1e-5))), det]
    force_valu = bool(rng.uniform() < 0.3)
    return d, model, force_valu, dict(n=n, separable=separable, curved=curved, kind=kind, shape=shape,
                                      rotation=rotation, force_valu=force_valu)


@pytest.mark.parametrize("seed", range(16))
def test_random_beamlets_vs_oracle(device, seed):
    d, model, force_valu, info = _random_case(seed)
    print(f"[random beamlets seed={seed} {info}]", flush=True)
    det = model[-1]
    autotune.clear()
    if force_valu:
        autotune.set_choice(autotune.BEAMLET_PATH, (len(d["x"]),) + tuple(det.shape), autotune.BEAMLET_PATH_VALU)
    try:
        got = make_gaussian_image(_gray(d), model).cpu().numpy()
    finally:
        autotune.clear()
    abcd = ray_matrix(model, Ray(0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    r2 = BO.grid_coords(det.pixel_size, det.shape, rotation=det.rotation)
    ref = BO.make_image(abcd=abcd, r2=r2, **d).reshape(det.shape)
    _check(got, ref)
