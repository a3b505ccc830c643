"""GPU parity of the Gaussian-beamlet kernels against the complex128 oracle.

Tolerance: |field - ref| <= 2e-5 * max|ref| everywhere (f32 recurrence and
accumulation), and phase within 1e-4 rad where |ref| > 1e-3 max|ref|
(compared modulo 2 pi)."""
import numpy as np
import pytest
import torch

from oracle import beamlet_oracle as BO
from temgymcore_amd.components import Biprism, Deflector, Detector, Lens, Scanner
from temgymcore_amd.gaussian import GaussianRay, evaluate_gaussian_input_image, make_gaussian_image
from temgymcore_amd.transfer import ray_matrix
from temgymcore_amd.ray import Ray

pytestmark = pytest.mark.gpu


def _check(got, ref, amp_tol=2e-5, phase_tol=1e-4):
    got = got.astype(np.complex128)
    scale = np.abs(ref).max()
    err = np.abs(got - ref)
    assert err.max() <= amp_tol * scale, f"max |d| {err.max() / scale:.3e} of max|ref|"
    strong = np.abs(ref) > 1e-3 * scale
    dphi = np.angle(got[strong] / ref[strong])
    assert np.abs(dphi).max() <= phase_tol, f"max phase error {np.abs(dphi).max():.3e} rad"


def _rays(n, rng, lam=2e-12, w=5e-8, spread=2e-7, tilt=1e-5, complex_amp=True, curved=True):
    x, y = rng.uniform(-spread, spread, (2, n))
    dx, dy = rng.uniform(-tilt, tilt, (2, n))
    amp = rng.uniform(0.5, 1.5, n) * (np.exp(1j * rng.uniform(-np.pi, np.pi, n)) if complex_amp else 1.0)
    waist = w * rng.uniform(0.7, 1.3, (n, 2))
    radii = rng.uniform(0.5, 2.0, (n, 2)) * (1 if curved else np.inf)
    if curved:
        radii[::3] = np.inf
    theta = rng.uniform(-1, 1, n)
    pl = rng.uniform(0, 1e-9, n)
    wl = np.full(n, lam)
    return dict(x=x, y=y, dx=dx, dy=dy, pathlength=pl, amplitude=amp, waist_xy=waist, radii=radii,
                wavelength=wl, theta=theta)


def _gray(d):
    n = len(d["x"])
    return GaussianRay(x=torch.from_numpy(d["x"]), y=torch.from_numpy(d["y"]), dx=torch.from_numpy(d["dx"]),
                       dy=torch.from_numpy(d["dy"]), z=torch.zeros(n, dtype=torch.float64),
                       pathlength=torch.from_numpy(d["pathlength"]), _one=torch.ones(n, dtype=torch.float64),
                       amplitude=torch.from_numpy(np.asarray(d["amplitude"])),
                       waist_xy=torch.from_numpy(d["waist_xy"]), radii_of_curv=torch.from_numpy(d["radii"]),
                       wavelength=torch.from_numpy(d["wavelength"]), theta=torch.from_numpy(d["theta"]))


@pytest.mark.parametrize("model_kind", ["free", "lens", "offsets"])
def test_make_gaussian_image_vs_oracle(device, model_kind):
    rng = np.random.default_rng(11)
    d = _rays(40, rng)
    if model_kind == "free":
        model = [Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(96, 80))]
    elif model_kind == "lens":
        model = [Lens(z=5e-4, focal_length=4e-4), Detector(z=1.4e-3, pixel_size=(4e-8, 4e-8), shape=(80, 96))]
    else:
        model = [Scanner(z=1e-4, scan_pos_x=3e-7, scan_pos_y=-2e-7), Deflector(z=2e-4, def_x=2e-5, def_y=-1e-5),
                 Detector(z=1e-3, pixel_size=(4e-8, 4e-8), shape=(72, 72), rotation=15.0)]
    got = make_gaussian_image(_gray(d), model).cpu().numpy()
    det = model[-1]
    abcd = ray_matrix(model, Ray(0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    r2 = BO.grid_coords(det.pixel_size, det.shape, rotation=det.rotation)
    ref = BO.make_image(abcd=abcd, r2=r2, **d).reshape(det.shape)
    _check(got, ref)


def test_evaluate_gaussian_input_image_vs_oracle(device):
    rng = np.random.default_rng(12)
    d = _rays(60, rng, lam=2.51e-12, w=2e-6, spread=6e-5, tilt=2e-3, curved=False)
    grid = Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(128, 128))
    got = evaluate_gaussian_input_image(_gray(d), grid).cpu().numpy()
    ref = BO.input_image(r1=BO.grid_coords(grid.pixel_size, grid.shape), **d).reshape(grid.shape)
    _check(got, ref)


def test_cutoff_drops_only_the_far_tails(device):
    rng = np.random.default_rng(13)
    d = _rays(200, rng, complex_amp=False, curved=False)
    model = [Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(128, 128))]
    full = make_gaussian_image(_gray(d), model).cpu().numpy().astype(np.complex128)
    cut = make_gaussian_image(_gray(d), model, cutoff=4.0).cpu().numpy().astype(np.complex128)
    # each dropped term is below exp(-16) of its beamlet's peak
    assert np.abs(full - cut).max() <= 200 * np.exp(-16) * 1.5 * np.abs(full).max() + 1e-6 * np.abs(full).max()
    assert not np.array_equal(full, cut)


def test_linearity_and_empty(device):
    rng = np.random.default_rng(14)
    d = _rays(300, rng)
    model = [Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(64, 200))]
    full = make_gaussian_image(_gray(d), model).cpu().numpy().astype(np.complex128)
    h1 = make_gaussian_image(_gray({k: v[:150] for k, v in d.items()}), model).cpu().numpy()
    h2 = make_gaussian_image(_gray({k: v[150:] for k, v in d.items()}), model).cpu().numpy()
    np.testing.assert_allclose(full, h1 + h2, rtol=0, atol=1e-5 * np.abs(full).max())
    e = make_gaussian_image(_gray({k: v[:0] for k, v in d.items()}), model)
    assert e.shape == (64, 200) and float(e.abs().max()) == 0.0


def aperture_workload(n=131072, seed=4, pix=None):
    """300 keV (lambda = 1.97 pm) beamlets uniform in a 10 um diameter aperture,
    waist 50 nm, propagated 0.3 m onto a 1024^2 detector."""
    rng = np.random.default_rng(seed)
    r = 5e-6 * np.sqrt(rng.uniform(0, 1, n))
    a = rng.uniform(0, 2 * np.pi, n)
    d = dict(x=r * np.cos(a), y=r * np.sin(a), dx=np.zeros(n), dy=np.zeros(n), pathlength=np.zeros(n),
             amplitude=np.ones(n), waist_xy=np.full((n, 2), 5e-8), radii=np.full((n, 2), np.inf),
             wavelength=np.full(n, 1.97e-12), theta=np.zeros(n))
    return d


def test_aperture_diffraction_full_size_sampled_pixels(device):
    """The aperture-diffraction workload at full size (131,072 beamlets, 0.3 m,
    1024^2 pixels of 40 nm, the 4-sigma cut) against the uncut complex128
    oracle: 224 pixels drawn inside the illuminated disc (radius 9 um: the
    5 um aperture plus about w(L) = 3.8 um) and 32 drawn over the whole
    detector.  Phases there reach ~1e3 rad, so the FP64 run starts and the
    double-double phase reduction are what the 1e-4 rad bar tests.

    Amplitude: |d| <= 2e-5 max|field| + tail(p), where tail(p) bounds what
    the cut drops at pixel p: every beamlet farther than 4 w(L) from p
    contributes at most (w0 / w(L)) exp(-16) (free space, untilted,
    round waist w0; w(L) = w0 sqrt(1 + (L / zR)^2)).  Inside the disc every
    beamlet is within 4 w(L), so tail = 0 there.  Phase: <= 1e-4 rad where
    |ref| > 1e-3 max and tail(p) < 1e-6 |ref|."""
    print("[beamlet aperture full size]", flush=True)
    d = aperture_workload()
    pix, L, w0, lam = 4e-8, 0.3, 5e-8, 1.97e-12
    det = Detector(z=L, pixel_size=(pix, pix), shape=(1024, 1024))
    got = make_gaussian_image(_gray(d), [det], cutoff=4.0).cpu().numpy()
    r2all = BO.grid_coords(det.pixel_size, det.shape)
    rng = np.random.default_rng(0)
    inner = np.nonzero(np.hypot(r2all[:, 0], r2all[:, 1]) < 9e-6)[0]
    sel = np.concatenate([rng.choice(inner, 224, replace=False), rng.choice(1024 * 1024, 32, replace=False)])
    r2 = r2all[sel]
    abcd = np.eye(5)
    abcd[0, 2] = abcd[1, 3] = L
    ref = BO.make_image(abcd=abcd, r2=r2, **d)
    g = got.reshape(-1)[sel].astype(np.complex128)
    peak = np.abs(got).max()
    wL = w0 * np.sqrt(1 + (L / (np.pi * w0 * w0 / lam)) ** 2)
    rho = np.hypot(r2[:, None, 0] - d["x"][None, :], r2[:, None, 1] - d["y"][None, :])
    tail = (rho > 3.99 * wL).sum(axis=1) * (w0 / wL) * np.exp(-16.0)
    assert np.all(tail[:224] == 0)
    err = np.abs(g - ref)
    assert np.all(err <= 2e-5 * peak + tail), f"max |d| - tail {np.max(err - tail) / peak:.3e} of max|field|"
    strong = (np.abs(ref) > 1e-3 * peak) & (tail < 1e-6 * np.abs(ref))
    assert strong.sum() >= 200
    dphi = np.abs(np.angle(g[strong] / ref[strong]))
    print(f"max |d| {err.max() / peak:.2e} of max|field|, max phase error {dphi.max():.2e} rad "
          f"on {strong.sum()} pixels", flush=True)
    assert dphi.max() <= 1e-4


def test_aperture_full_field_both_tile_sums_agree(device):
    """The whole 1024^2 field of the aperture workload from the matrix-core
    tile sum (default for these separable beamlets) and from the VALU
    recurrence (forced): they differ by rounding and by the sub-cut terms the
    matrix path keeps inside listed tiles, at most 131,072 (w0 / w(L)) e^-16
    in field units, far below the 2e-5 max|field| bound.  Phase within
    1e-4 rad wherever |field| > 1e-2 max (two f32 evaluations, each good to
    ~1e-7 max|field| absolute, can differ by ~1e-4 relative where the field
    is only 1e-3 of its maximum; the oracle test bounds each path there)."""
    from temgymcore_amd import autotune

    print("[beamlet aperture full field, matrix vs recurrence]", flush=True)
    d = aperture_workload()
    n = len(d["x"])
    det = Detector(z=0.3, pixel_size=(4e-8, 4e-8), shape=(1024, 1024))
    autotune.clear()
    try:
        mat = make_gaussian_image(_gray(d), [det], cutoff=4.0).cpu().numpy().astype(np.complex128)
        autotune.set_choice(autotune.BEAMLET_PATH, (n, 1024, 1024), autotune.BEAMLET_PATH_VALU)
        rec = make_gaussian_image(_gray(d), [det], cutoff=4.0).cpu().numpy().astype(np.complex128)
    finally:
        autotune.clear()
    peak = np.abs(rec).max()
    err = np.abs(mat - rec)
    strong = np.abs(rec) > 1e-2 * peak
    dphi = np.abs(np.angle(mat[strong] / rec[strong]))
    mid = (np.abs(rec) > 1e-3 * peak) & ~strong
    rel_mid = err[mid] / np.abs(rec[mid])
    print(f"max |d| {err.max() / peak:.2e} of max|field|; max phase difference {dphi.max():.2e} rad on "
          f"{strong.sum()} pixels above 1e-2 max; relative difference on {mid.sum()} pixels in "
          f"(1e-3, 1e-2] max: median {np.median(rel_mid):.2e}, max {rel_mid.max():.2e}", flush=True)
    assert err.max() <= 2e-5 * peak
    assert strong.sum() > 10000 and dphi.max() <= 1e-4
    # dim pixels: the 2e-5 max|field| bound allows 2e-2 relative at 1e-3 max;
    # measured ~1e-4, so 1e-3 still catches a systematic error between paths
    assert mid.sum() > 10000 and rel_mid.max() <= 1e-3


def _oracle_abcd(model_t, x, y, dx, dy, h=1e-9):
    """Per-beamlet 5x5 Jacobians of the oracle model by central differences
    (exact for the piecewise-affine elements away from their kinks)."""
    from oracle import ray_oracle as O

    n = len(x)
    base = np.stack([x, y, dx, dy, np.ones(n)], 1).astype(np.float64)
    J = np.zeros((n, 5, 5))
    for j in range(5):
        outs = []
        for sgn in (1, -1):
            v = base.copy()
            v[:, j] += sgn * h
            r = O.run_to_end(O.make_ray(v[:, 0], v[:, 1], v[:, 2], v[:, 3], 0.0, 0.0, v[:, 4]), model_t)
            outs.append(np.stack([r["x"], r["y"], r["dx"], r["dy"], np.broadcast_to(r["one"], (n,))], 1))
        J[:, :, j] = (outs[0] - outs[1]) / (2 * h)
    return J


def test_biprism_beamlet_image_vs_oracle(device):
    """A Biprism makes the ray-transfer matrix depend on the beamlet: each
    beamlet gets its own from forward-mode tangents on the GPU."""
    rng = np.random.default_rng(31)
    n = 24
    d = _rays(n, rng, tilt=1e-6)
    # keep every central ray clear of the biprism edge (x = 0 at its plane)
    d["x"] = np.sign(rng.uniform(-1, 1, n)) * rng.uniform(4e-8, 2e-7, n)
    model = [Biprism(z=1e-4, def_x=2e-5), Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(72, 80))]
    got = make_gaussian_image(_gray(d), model).cpu().numpy()
    J = _oracle_abcd([("biprism", 1e-4, 2e-5), ("detector", 1e-3)], d["x"], d["y"], d["dx"], d["dy"])
    assert not np.allclose(J[0, 2, 4], J[np.argmax(np.sign(d["x"]) != np.sign(d["x"][0])), 2, 4])  # two sides
    r2 = BO.grid_coords((3e-8, 3e-8), (72, 80))
    ref = BO.make_image(d["x"], d["y"], d["dx"], d["dy"], d["pathlength"], d["amplitude"], d["waist_xy"], d["radii"],
                        d["wavelength"], d["theta"], J, r2).reshape(72, 80)
    _check(got, ref)


def test_traced_abcd_matches_host_matrices(device):
    from temgymcore_amd.gaussian import _abcd_traced
    from temgymcore_amd.transfer import ray_matrix

    rng = np.random.default_rng(32)
    d = _rays(6, rng)
    model = [Lens(z=2e-4, focal_length=5e-4), Biprism(z=4e-4, def_x=1e-5),
             Deflector(z=6e-4, def_x=1e-6, def_y=-2e-6), Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(8, 8))]
    J = _abcd_traced(_gray(d), model, device, 6).cpu().numpy()
    for b in range(6):
        ray = Ray(float(d["x"][b]), float(d["y"][b]), float(d["dx"][b]), float(d["dy"][b]), 0.0, 0.0)
        np.testing.assert_allclose(J[b], np.asarray(ray_matrix(model, ray)), rtol=1e-12, atol=1e-15)


def _separable(d):
    """theta = 0: astigmatic waists and curvatures stay along the axes, so an
    axis-aligned model leaves Phi without a px*py term."""
    return dict(d, theta=np.zeros(len(d["x"])))


@pytest.mark.parametrize("path", ["auto", "valu"])
@pytest.mark.parametrize("model_kind", ["free", "lens"])
def test_separable_beamlets_both_tile_sums_vs_oracle(device, path, model_kind):
    """Separable beamlets (tilts, offsets, complex amplitudes, astigmatic
    waists and curvatures along the axes) take the matrix-core tile sum by
    default; the VALU recurrence, forced through the tuning knob, must give
    the same field.  150 x 200 pixels: partial edge tiles on both axes."""
    from temgymcore_amd import autotune

    rng = np.random.default_rng(21)
    d = _separable(_rays(40, rng))
    det = Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(150, 200))
    model = [det] if model_kind == "free" else [Lens(z=4e-4, focal_length=5e-4), det]
    autotune.clear()
    if path == "valu":
        autotune.set_choice(autotune.BEAMLET_PATH, (40, 150, 200), autotune.BEAMLET_PATH_VALU)
    try:
        got = make_gaussian_image(_gray(d), model).cpu().numpy()
    finally:
        autotune.clear()
    abcd = ray_matrix(model, Ray(0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    ref = BO.make_image(abcd=abcd, r2=BO.grid_coords(det.pixel_size, det.shape), **d).reshape(det.shape)
    _check(got, ref)


def test_one_non_separable_beamlet_selects_the_recurrence(device):
    """A rotated astigmatic beamlet has a px*py term: the whole launch must
    take the VALU recurrence (the matrix-core product would drop that term)."""
    rng = np.random.default_rng(22)
    d = _separable(_rays(30, rng))
    d["theta"][7] = 0.6
    assert d["waist_xy"][7, 0] != d["waist_xy"][7, 1]
    det = Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(96, 128))
    got = make_gaussian_image(_gray(d), [det]).cpu().numpy()
    abcd = ray_matrix([det], Ray(0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    ref = BO.make_image(abcd=abcd, r2=BO.grid_coords(det.pixel_size, det.shape), **d).reshape(det.shape)
    _check(got, ref)


def test_separable_input_plane_image_vs_oracle(device):
    rng = np.random.default_rng(23)
    d = _separable(_rays(60, rng, lam=2.51e-12, w=2e-6, spread=6e-5, tilt=2e-3, curved=True))
    grid = Detector(z=0.0, pixel_size=(1e-6, 1e-6), shape=(130, 100))
    got = evaluate_gaussian_input_image(_gray(d), grid).cpu().numpy()
    ref = BO.input_image(r1=BO.grid_coords(grid.pixel_size, grid.shape), **d).reshape(grid.shape)
    _check(got, ref)


def _single(x0, y0, dx, dy, lam=500e-9, w0=1e-4):
    one = lambda v: np.array([v], dtype=np.float64)  # noqa: E731
    return dict(x=one(x0), y=one(y0), dx=one(dx), dy=one(dy), pathlength=one(0.0), amplitude=one(1.0),
                waist_xy=np.array([[w0, w0]]), radii=np.full((1, 2), np.inf), wavelength=one(lam), theta=one(0.0))


def test_tilted_beam_lands_in_its_quadrant(device):
    """Reference test_gaussians.py:684-733: a beam tilted by (dx, dy) peaks on
    the detector 1 m away on the side of its tilt, by dx and dy metres (10
    pixels here, to within the argmax's pixel)."""
    det = Detector(z=1.0, pixel_size=(1e-4, 1e-4), shape=(256, 256))
    xs, ys = det.coords_1d
    Y, X = np.meshgrid(np.asarray(ys, dtype=np.float64), np.asarray(xs, dtype=np.float64), indexing="ij")

    def peak(img):
        my, mx = np.unravel_index(np.argmax(np.abs(img)), img.shape)
        return X[my, mx], Y[my, mx]

    xc, yc = peak(evaluate_gaussian_input_image(_gray(_single(0.0, 0.0, 0.0, 0.0)), det).cpu().numpy())
    for dx, dy in ((1e-3, 1e-3), (1e-3, -1e-3), (-1e-3, 1e-3), (-1e-3, -1e-3)):
        xm, ym = peak(make_gaussian_image(_gray(_single(0.0, 0.0, dx, dy)), [det]).cpu().numpy())
        assert np.sign(xm - xc) == np.sign(dx) and np.sign(ym - yc) == np.sign(dy), (dx, dy, xm, ym)
        assert abs((xm - xc) - dx) <= 1.5e-4 and abs((ym - yc) - dy) <= 1.5e-4, (dx, dy, xm, ym)
