"""Pins the ray-branch oracle against the reference's own golden values,
known-answer tables and closed forms, plus brute force for binning."""
import math

import numpy as np
import pytest

from oracle import ray_oracle as O


def test_readme_golden_single_ray():
    # reference README.md:35-52: Ray(0.1, 0.2, 0.3, 0.4, 0, 0) -> Lens(0.5, f=1) -> Detector(1.0)
    r = O.run_to_end(O.make_ray(0.1, 0.2, 0.3, 0.4), [("lens", 0.5, 1.0), ("detector", 1.0)])
    assert r["x"] == pytest.approx(0.275, abs=1e-15)
    assert r["y"] == pytest.approx(0.4, abs=1e-15)
    assert r["dx"] == pytest.approx(0.05, abs=1e-15)
    assert r["dy"] == pytest.approx(0.0, abs=1e-15)
    assert r["z"] == 1.0
    # printed 0.89 in the README; by hand 0.5 - 0.25/2 + 0.5 + 0.5*... = 0.88875
    assert r["pathlength"] == pytest.approx(0.88875, abs=1e-15)


def _fd_matrix(model, r0):
    """5x5 ABCD by central differences (x, y, dx, dy, one)."""
    keys = ("x", "y", "dx", "dy", "one")
    J = np.zeros((5, 5))
    for j, k in enumerate(keys):
        h = 1e-6
        a = dict(r0)
        b = dict(r0)
        a[k] = r0[k] + h
        b[k] = r0[k] - h
        ra, rb = O.run_to_end(a, model), O.run_to_end(b, model)
        for i, kk in enumerate(keys):
            J[i, j] = (ra[kk] - rb[kk]) / (2 * h)
    return J


def test_readme_abcd_matrix():
    # reference README.md:230-234
    expected = np.array([
        [0.5, 0, 0.75, 0, 0], [0, 0.5, 0, 0.75, 0], [-1, 0, 0.5, 0, 0], [0, -1, 0, 0.5, 0], [0, 0, 0, 0, 1]])
    J = _fd_matrix([("lens", 0.5, 1.0), ("detector", 1.0)], O.make_ray(0.1, 0.2, 0.3, 0.4))
    np.testing.assert_allclose(J, expected, atol=1e-8)


def test_readme_parameter_gradient():
    # reference README.md:124-136: d x_out / d (f, z) = (0.125, 0.1)
    def x_out(f, z):
        return O.run_to_end(O.make_ray(0.1, 0.2, 0.3, 0.4), [("lens", z, f), ("detector", 1.0)])["x"]

    h = 1e-6
    assert (x_out(1 + h, 0.5) - x_out(1 - h, 0.5)) / (2 * h) == pytest.approx(0.125, abs=1e-8)
    assert (x_out(1, 0.5 + h) - x_out(1, 0.5 - h)) / (2 * h) == pytest.approx(0.1, abs=1e-8)


@pytest.mark.parametrize("seed", range(3))
def test_propagate_closed_form(seed):
    # reference test_rays.py:26-39
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.1, 10)
    x0, y0, dx0, dy0, z0, pl0 = rng.uniform(-10, 10, 6)
    r = O.propagate(O.make_ray(x0, y0, dx0, dy0, z0, pl0), d)
    assert r["x"] == pytest.approx(x0 + dx0 * d, abs=1e-12)
    assert r["y"] == pytest.approx(y0 + dy0 * d, abs=1e-12)
    assert r["z"] == pytest.approx(z0 + d, abs=1e-12)
    assert r["pathlength"] == pytest.approx(pl0 + d, abs=1e-12)


def test_propagation_matrix():
    # reference test_rays.py:82-103: Jacobian of propagation = free-space 5x5
    d = 2.5
    T = np.eye(5)
    T[0, 2] = T[1, 3] = d
    r0 = O.make_ray(0.5, -0.5, 0.1, -0.2, 1.0)
    J = _fd_matrix([("plane", 1.0 + d)], r0)
    np.testing.assert_allclose(J, T, atol=1e-8)


def test_scanner_and_descanner_closed_form():
    # reference test_component.py:186-259
    rng = np.random.default_rng(11)
    spx, spy = rng.uniform(-5, 5, 2)
    stx, sty = rng.uniform(-0.5, 0.5, 2)
    x, y, dx, dy = rng.uniform(-5, 5, 4)
    e = rng.random(12)
    r = O.scanner(O.make_ray(x, y, dx, dy), spx, spy, stx, sty)
    assert (r["x"], r["y"], r["dx"], r["dy"]) == pytest.approx((x + spx, y + spy, dx + stx, dy + sty))
    r = O.descanner(O.make_ray(x, y, dx, dy), spx, spy, stx, sty, e)
    assert r["x"] == pytest.approx(x + spx * e[0] + spy * e[1] + e[8] - spx)
    assert r["y"] == pytest.approx(y + spx * e[2] + spy * e[3] + e[9] - spy)
    assert r["dx"] == pytest.approx(dx + spx * e[4] + spy * e[5] + e[10] - stx)
    assert r["dy"] == pytest.approx(dy + spx * e[6] + spy * e[7] + e[11] - sty)


def test_descanner_matrix_column():
    # reference test_component.py:313-353: 5th column K1..K4, identity elsewhere
    e = np.random.default_rng(3).random(12)
    spx, spy = 1.5, -2.0
    J = _fd_matrix([("descanner", 0.0, spx, spy, 0.0, 0.0, e)], O.make_ray(0.0, 0.0, 0.0, 0.0))
    T = np.eye(5)
    T[0, 4] = spx * e[0] + spy * e[1] + e[8] - spx
    T[1, 4] = spx * e[2] + spy * e[3] + e[9] - spy
    T[2, 4] = spx * e[4] + spy * e[5] + e[10]
    T[3, 4] = spx * e[6] + spy * e[7] + e[11]
    np.testing.assert_allclose(J, T, atol=1e-7)


def test_lens_matrix_and_pathlength():
    # thin lens ABCD (tests/transfer_matrices.py lens matrix) and the paraxial phase
    f = 0.7
    J = _fd_matrix([("lens", 0.0, f)], O.make_ray(0.01, 0.02, 0.0, 0.0))
    T = np.eye(5)
    T[2, 0] = T[3, 1] = -1 / f
    np.testing.assert_allclose(J, T, atol=1e-8)
    r = O.lens(O.make_ray(0.3, 0.4, 0.0, 0.0), f)
    assert r["pathlength"] == pytest.approx(-(0.25) / (2 * f))


def test_rotator_and_deflector():
    r = O.rotator(O.make_ray(1.0, 0.0, 0.0, 1.0), 90.0)
    assert (r["x"], r["y"], r["dx"], r["dy"]) == pytest.approx((0.0, 1.0, -1.0, 0.0), abs=1e-15)
    r = O.deflector(O.make_ray(2.0, 3.0, 0.1, 0.2, pathlength=1.0), 0.5, -0.5)
    assert (r["dx"], r["dy"]) == pytest.approx((0.6, -0.3))
    assert r["pathlength"] == pytest.approx(1.0 + 0.1 * 2.0 + 0.2 * 3.0)  # incoming slopes


TABLE_M2P = [
    ((0.0, 0.0), 0.0, (5, 5)), ((-0.5, 0.5), 0.0, (0, 0)), ((0.5, -0.5), 0.0, (10, 10)),
    ((0.0, 0.5), 0.0, (0, 5)), ((-0.5, 0.0), 0.0, (5, 0)),
    ((0.0, 0.0), 90.0, (5, 5)), ((-0.5, 0.5), 90.0, (10, 0)), ((0.5, -0.5), 90.0, (0, 10)),
    ((0.0, 0.5), 90.0, (5, 0)), ((-0.5, 0.0), 90.0, (10, 5)),
]


@pytest.mark.parametrize("xy, rot, expected", TABLE_M2P)
def test_metres_to_pixels_table(xy, rot, expected):
    # reference test_component.py:73-99 (ScanGrid) and :131-150 (Detector, rot 0)
    Tinv = O.metres_to_pixels_matrix((0.0, 0.0), (0.1, 0.1), (11, 11), rotation=rot)
    py, px = O.metres_to_pixels(xy[0], xy[1], Tinv)
    assert (int(py), int(px)) == expected


@pytest.mark.parametrize("xy, rot, px", TABLE_M2P)
def test_pixels_to_metres_table(xy, rot, px):
    # reference test_component.py:102-128, :153-174
    T = O.pixels_to_metres_matrix((0.0, 0.0), (0.1, 0.1), (11, 11), rotation=rot)
    my, mx = O.transform(px[0], px[1], T)
    np.testing.assert_allclose((mx, my), xy, atol=1e-12)


@pytest.mark.parametrize("rot", [-137.0, -12.0, 12.0, 75.0, 170.0])
def test_scan_step_vector(rot):
    # reference test_component.py:356-378: one x-pixel step = (cos t * s, -sin t * s)
    T = O.pixels_to_metres_matrix((0.0, 0.0), (0.1, 0.1), (11, 11), rotation=rot)
    y0, x0 = O.transform(5, 5, T)
    y1, x1 = O.transform(5, 6, T)
    t = math.radians(rot)
    np.testing.assert_allclose((x1 - x0, y1 - y0), (0.1 * math.cos(t), -0.1 * math.sin(t)), atol=1e-12)


def test_round_half_even_and_cast():
    assert list(O.round_to_i32([0.5, 1.5, 2.5, -0.5, -1.5])) == [0, 2, 2, 0, -2]
    assert list(O.round_to_i32([np.nan, 1e30, -1e30])) == [0, 2147483647, -2147483648]


def test_binning_brute_force():
    rng = np.random.default_rng(5)
    py = rng.integers(-3, 12, 4000)
    px = rng.integers(-3, 12, 4000)
    a = O.into_image(py, px, (9, 10))
    b = O.into_image_loop(py, px, (9, 10))
    np.testing.assert_array_equal(a, b)
    inside = ((py >= 0) & (py < 9) & (px >= 0) & (px < 10)).sum()
    assert a.sum() == inside


def test_plain_workload_histogram_consistency():
    x, y, dx, dy = O.plain_workload(20000, seed=0)
    r, hist = O.plain_run(x, y, dx, dy)
    assert r["x"].dtype == np.float32 and r["pathlength"].dtype == np.float32
    # x_det = 0.5 x + 0.75 dx (README ABCD): f32 result within rounding of the f64 value
    ref = 0.5 * x.astype(np.float64) + 0.75 * dx.astype(np.float64)
    np.testing.assert_allclose(r["x"], ref, rtol=0, atol=4e-8)
    assert hist.sum() == x.size  # every ray lands inside the 128x128 detector
    # brute force over the same pixel indices
    Tinv = O.metres_to_pixels_matrix(**O.PLAIN_DET)
    py, px = O.metres_to_pixels(r["x"], r["y"], Tinv)
    np.testing.assert_array_equal(hist, O.into_image_loop(py, px, (128, 128)))


def test_planted_rounding_mistake_is_caught():
    # floor instead of half-even rounding must break the known-answer table
    Tinv = O.metres_to_pixels_matrix((0.0, 0.0), (0.1, 0.1), (11, 11))
    py, px = O.transform(0.5, -0.5, Tinv)  # (x, y) = (-0.5, 0.5)? no: y=0.5, x=-0.5
    good = (int(O.round_to_i32(py)), int(O.round_to_i32(px)))
    bad = (int(np.floor(py - 1e-9)), int(np.floor(px - 1e-9)))
    assert good == (0, 0) and bad != good


# --------------------------------------------------------------------------
# AberratedLensKrivanek (reference components.py:177-215, aberrations.py)
# --------------------------------------------------------------------------
def _golden_aberrated():
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "aberrated_probe_rays.json")) as fh:
        return json.load(fh)


def test_krivanek_notebook_golden_rays():
    """The reference's printed run_to_end of examples/aberrated_probe.ipynb
    (cell 9): spherical aberration C30 = 1e8 m at f = 0.25 mm.  The edge
    rays' detector positions are pure aberration (x = -dW/dax)."""
    g = _golden_aberrated()
    st = g["setup"]
    x, y = O.fibonacci_spiral(st["n_beams"], st["radius"], alpha=st["alpha"])
    n, F1 = len(x), st["F1"]
    r = O.make_ray(x, y, np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n))
    o = O.run_to_end(r, [("plane", 0.0), ("kriv", F1, F1, st["coeffs"]), ("detector", 2 * F1)])
    idx = np.array(g["index"])
    for k in ("x", "y", "dx", "dy"):
        np.testing.assert_allclose(o[k][idx], g[k], rtol=6e-9, atol=0, err_msg=k)
    np.testing.assert_allclose(o["pathlength"][idx], g["pathlength"], rtol=1e-8)


def _kriv_random(rng, scale=1.0):
    c = {}
    for k in O.KRIVANEK_FIELDS:
        c[k] = rng.uniform(-np.pi, np.pi) if k.startswith("phi") else scale * rng.uniform(-1, 1)
    return c


def test_krivanek_gradient_vs_central_differences():
    """grad_W (polar derivatives mapped to slopes) against Richardson central
    differences of W itself, every coefficient non-zero."""
    rng = np.random.default_rng(41)
    c = _kriv_random(rng)
    ax, ay = rng.uniform(-0.6, 0.6, 50), rng.uniform(-0.6, 0.6, 50)

    def W(u, v):
        return O.krivanek_W(np.hypot(u, v), np.arctan2(v, u), c)

    def cd(f, h):
        return (f(h) - f(-h)) / (2 * h)

    h = 1e-4
    gx_fd = (4 * cd(lambda e: W(ax + e, ay), h / 2) - cd(lambda e: W(ax + e, ay), h)) / 3
    gy_fd = (4 * cd(lambda e: W(ax, ay + e), h / 2) - cd(lambda e: W(ax, ay + e), h)) / 3
    gx, gy = O.krivanek_grad(ax, ay, c)
    np.testing.assert_allclose(gx, gx_fd, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(gy, gy_fd, rtol=1e-8, atol=1e-10)


def test_krivanek_closed_forms():
    """Defocus C10: W = C10 a^2/2, grad = C10 (ax, ay); spherical C30:
    W = C30 a^4/4, grad = C30 a^2 (ax, ay); two-fold astigmatism C12 at
    phi12: W = C12/2 a^2 cos(2(phi - phi12))."""
    ax, ay = np.array([0.3, -0.1, 0.0]), np.array([0.2, 0.4, 0.0])
    a2 = ax * ax + ay * ay
    gx, gy = O.krivanek_grad(ax, ay, {"C10": 2.0})
    np.testing.assert_allclose(gx, 2.0 * ax, rtol=1e-14)
    np.testing.assert_allclose(gy, 2.0 * ay, rtol=1e-14)
    gx, gy = O.krivanek_grad(ax, ay, {"C30": 3.0})
    np.testing.assert_allclose(gx, 3.0 * a2 * ax, rtol=1e-14)
    np.testing.assert_allclose(gy, 3.0 * a2 * ay, rtol=1e-14)
    a, ph = np.sqrt(a2), np.arctan2(ay, ax)
    np.testing.assert_allclose(O.krivanek_W(a, ph, {"C30": 3.0}), 0.75 * a2 * a2, rtol=1e-14)
    np.testing.assert_allclose(O.krivanek_W(a, ph, {"C12": 2.0, "phi12": 0.4}),
                               a2 * np.cos(2 * (ph - 0.4)), rtol=1e-14)
    # the ray update: defocus C10 scales the ideal slope by (1 - C10/f)
    r = O.aberrated_lens(O.make_ray(0.01, -0.02, 0.003, 0.001), 0.5, {"C10": 0.1})
    assert r["dx"] == pytest.approx((-0.01 / 0.5 + 0.003) * (1 - 0.1 / 0.5), rel=1e-14)


def test_krivanek_rotation_covariance():
    """Rotating the slopes by t and every azimuth by t leaves W unchanged and
    rotates its gradient."""
    rng = np.random.default_rng(42)
    c = _kriv_random(rng)
    t = 0.7
    cr = {k: (v + t if k.startswith("phi") else v) for k, v in c.items()}
    ax, ay = rng.uniform(-0.5, 0.5, 20), rng.uniform(-0.5, 0.5, 20)
    bx, by = np.cos(t) * ax - np.sin(t) * ay, np.sin(t) * ax + np.cos(t) * ay
    W0 = O.krivanek_W(np.hypot(ax, ay), np.arctan2(ay, ax), c)
    W1 = O.krivanek_W(np.hypot(bx, by), np.arctan2(by, bx), cr)
    np.testing.assert_allclose(W1, W0, rtol=1e-12, atol=1e-14)
    gx, gy = O.krivanek_grad(ax, ay, c)
    hx, hy = O.krivanek_grad(bx, by, cr)
    np.testing.assert_allclose(hx, np.cos(t) * gx - np.sin(t) * gy, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(hy, np.sin(t) * gx + np.cos(t) * gy, rtol=1e-11, atol=1e-13)


def test_krivanek_planted_mistake_is_caught():
    """The central-difference check fails a gradient whose azimuthal part
    has the wrong sign (a plausible slip in the dW/dphi mapping)."""
    rng = np.random.default_rng(43)
    c = {"C23": 1.0, "phi23": 0.3}
    ax, ay = rng.uniform(-0.5, 0.5, 10), rng.uniform(-0.5, 0.5, 10)
    h = 1e-5
    fd = (O.krivanek_W(np.hypot(ax + h, ay), np.arctan2(ay, ax + h), c)
          - O.krivanek_W(np.hypot(ax - h, ay), np.arctan2(ay, ax - h), c)) / (2 * h)
    gx, gy = O.krivanek_grad(ax, ay, c)
    assert np.allclose(gx, fd, rtol=1e-6, atol=1e-9)
    a = np.hypot(ax, ay)
    dW_da, dW_dphi = (gx * ax + gy * ay) / a, gy * ax - gx * ay
    gx_bad = dW_da * ax / a + dW_dphi * ay / a ** 2
    assert not np.allclose(gx_bad, fd, rtol=1e-6, atol=1e-9)


# ---- Biprism (reference components.py:526-559) ----------------------------------
#
# Pins of ray_oracle.biprism: the reference's tests test_component.py:398-465
# (the deflection column of the Jacobian for x > 0 and x < 0, the deflection
# grown over the propagation to the detector, and the lens + biprism system
# against the product of the analytic 5x5 matrices of tests/transfer_matrices.py
# written out below), a hand-computed pathlength (incoming slopes), and the
# identity with a Deflector of +-def_x away from the apex.  Planted mistakes
# (sign(-x), outgoing slopes in the pathlength) must fail.

def _prop5(z):
    m = np.eye(5)
    m[0, 2] = m[1, 3] = z
    return m


def _lens5(f):
    m = np.eye(5)
    m[2, 0] = m[3, 1] = -1.0 / f
    return m


def _biprism5(d):
    m = np.eye(5)
    m[2, 4] = d
    return m


def _fd_matrix_affine(model, r0, h):
    """5x5 Jacobian of a model that is affine near r0 (no kink within +-h)."""
    keys = ("x", "y", "dx", "dy", "one")
    J = np.zeros((5, 5))
    for j, k in enumerate(keys):
        a, b = dict(r0), dict(r0)
        a[k] = r0[k] + h
        b[k] = r0[k] - h
        ra, rb = O.run_to_end(a, model), O.run_to_end(b, model)
        for i, kk in enumerate(keys):
            J[i, j] = (ra[kk] - rb[kk]) / (2 * h)
    return J


def _check_biprism_deflection_column():
    # test_component.py:398-407: d(dx)/d(one) = +def_x just right of the apex;
    # and -def_x just left of it
    d = 1e-3
    for x0, sgn in ((1e-15, 1.0), (-1e-15, -1.0)):
        J = _fd_matrix_affine([("biprism", 0.0, d)], O.make_ray(x0, 0.0, 0.0, 0.0), 1e-6)
        assert abs(J[2, 4] - sgn * d) < 1e-12
        assert abs(J[3, 4]) < 1e-15 and abs(J[0, 4]) < 1e-15


def _check_biprism_with_propagation():
    # test_component.py:410-424: d x_det / d one = def_x (z_det - z_biprism)
    d, z_det = 1e-3, 0.234
    for x0, sgn in ((1e-15, 1.0), (-1e-15, -1.0)):
        J = _fd_matrix_affine([("biprism", 0.0, d), ("detector", z_det)], O.make_ray(x0, 0.0, 0.0, 0.0), 1e-6)
        assert abs(J[0, 4] - sgn * d * z_det) < 1e-12
        assert abs(J[2, 4] - sgn * d) < 1e-12


def _check_biprism_lens_system():
    # test_component.py:426-465: source plane at the defocus, lens, biprism
    # half-way to the detector; ABCD = P(z3) Bp(d) P(z2) L(f) P(z1).  The
    # system magnifies by -4.5 up to the biprism, so a central ray just left
    # of the axis at the source crosses the biprism right of its apex; the
    # Jacobian is taken inside that affine region (source x = -1e-3)
    M1, F1, defocus, d = -10, 0.0002, 1e-4, 1e-4
    zl = abs(F1 * (1 / M1 - 1))
    z2_full = abs(F1 * (1 - M1))
    zb, zd = zl + z2_full / 2, zl + z2_full
    model = [("plane", defocus), ("lens", zl, F1), ("biprism", zb, d), ("detector", zd)]
    r0 = O.make_ray(-1e-3, 0.0, 0.0, 0.0, z=defocus)
    J = _fd_matrix_affine(model, r0, 1e-5)
    exp = _prop5(zd - zb) @ _biprism5(d) @ _prop5(zb - zl) @ _lens5(F1) @ _prop5(zl - defocus)
    np.testing.assert_allclose(J, exp, rtol=1e-8, atol=1e-9)
    # the central ray itself (-1e-15) is on the same side at the biprism
    mid = O.run_steps(O.make_ray(-1e-15, 0.0, 0.0, 0.0, z=defocus), model)[4]
    assert mid["x"] > 0


def _check_biprism_pathlength_and_deflector_identity():
    # pathlength gains dx_in * x + dy_in * y (incoming slopes), by hand:
    # 0.01 * 0.2 + 0.02 * 0.3 = 0.008
    r = O.apply(O.make_ray(0.2, 0.3, 0.01, 0.02, pathlength=0.5), ("biprism", 0.0, 1e-3))
    assert r["pathlength"] == pytest.approx(0.508, abs=1e-15)
    assert r["dx"] == pytest.approx(0.011, abs=1e-15) and r["dy"] == 0.02
    # away from the apex the biprism is a Deflector(+-def_x, 0)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, 64)
    x[np.abs(x) < 1e-3] = 0.5
    ray = O.make_ray(x, rng.uniform(-1, 1, 64), rng.uniform(-1e-2, 1e-2, 64), rng.uniform(-1e-2, 1e-2, 64),
                     pathlength=rng.uniform(0, 1, 64))
    got = O.apply(ray, ("biprism", 0.0, 2e-3))
    right = O.apply(ray, ("deflector", 0.0, 2e-3, 0.0))
    left = O.apply(ray, ("deflector", 0.0, -2e-3, 0.0))
    for k in ("x", "y", "dx", "dy", "pathlength"):
        np.testing.assert_array_equal(got[k], np.where(x > 0, right[k], left[k]))


_BIPRISM_CHECKS = (_check_biprism_deflection_column, _check_biprism_with_propagation, _check_biprism_lens_system,
                   _check_biprism_pathlength_and_deflector_identity)


@pytest.mark.parametrize("check", _BIPRISM_CHECKS, ids=lambda f: f.__name__)
def test_biprism_pins(check):
    check()


def _planted_biprism_flipped_side(r, def_x):
    return dict(r, dx=r["dx"] + def_x * r["one"] * np.sign(-r["x"]),
                pathlength=r["pathlength"] + r["dx"] * r["x"] + r["dy"] * r["y"])


def _planted_biprism_outgoing_slopes(r, def_x):
    dx = r["dx"] + def_x * r["one"] * np.sign(r["x"])
    return dict(r, dx=dx, pathlength=r["pathlength"] + dx * r["x"] + r["dy"] * r["y"])


@pytest.mark.parametrize("plant", [_planted_biprism_flipped_side, _planted_biprism_outgoing_slopes],
                         ids=lambda f: f.__name__)
def test_biprism_planted_mistakes_fail(plant, monkeypatch):
    monkeypatch.setattr(O, "biprism", plant)
    caught = 0
    for check in _BIPRISM_CHECKS:
        try:
            check()
        except AssertionError:
            caught += 1
    assert caught >= 1, "a planted mistake in the biprism passed every pin"


def test_thick_lens_closed_form():
    """ThickLens(z_po=0.5, z_pi=0.6, f=1) then a detector at 1.0, on the README
    ray: thin lens at 0.5 (x = 0.25, dx = 0.3 - 0.25 = 0.05), the ray moves to
    0.6, then 0.4 m of drift: x = 0.27, y = 0.4 + 0 * 0.4 = 0.4; pathlength
    0.5 - (0.25^2 + 0.4^2) / 2 + 0.4 = 0.78875.  Per axis the ABCD is
    drift(0.4) lens(1) drift(0.5) = [[0.6, 0.7], [-1, 0.5]].  (Moving z the
    wrong way would end at x = 0.28.)"""
    model = [("thicklens", 0.5, 0.6, 1.0), ("detector", 1.0)]
    r = O.run_to_end(O.make_ray(0.1, 0.2, 0.3, 0.4), model)
    assert r["x"] == pytest.approx(0.27, abs=1e-15)
    assert r["y"] == pytest.approx(0.4, abs=1e-15)
    assert r["dx"] == pytest.approx(0.05, abs=1e-15)
    assert r["dy"] == pytest.approx(0.0, abs=1e-15)
    assert r["z"] == pytest.approx(1.0, abs=1e-15)
    assert r["pathlength"] == pytest.approx(0.78875, abs=1e-15)
    J = _fd_matrix(model, O.make_ray(0.0, 0.0, 0.0, 0.0))
    expect = np.array([[0.6, 0, 0.7, 0, 0], [0, 0.6, 0, 0.7, 0], [-1, 0, 0.5, 0, 0], [0, -1, 0, 0.5, 0],
                       [0, 0, 0, 0, 1.0]])
    np.testing.assert_allclose(J, expect, atol=1e-8)
