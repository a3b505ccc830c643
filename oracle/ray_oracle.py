"""Ray-branch oracle (numpy) -- test infrastructure, see oracle/__init__.py.

A ray is a dict with keys x, y, dx, dy, z, pathlength, one.  Values are numpy
arrays (per ray) or Python floats (uniform).  numpy >= 2 follows NEP 50: a
Python float meeting an f32 array is rounded to f32 and the operation runs in
f32 -- the same "weak scalar" rule JAX applies to the reference, so writing
the reference formulas with numpy arrays reproduces its rounding: every
elementwise numpy operation is one IEEE operation, without FMA contraction.

A model is a list of (kind, params) tuples:
  ("plane", z) ("lens", z, f) ("detector", z) ("scanner", z, spx, spy, stx, sty)
  ("descanner", z, spx, spy, stx, sty, err12) ("deflector", z, defx, defy)
  ("rotator", z, angle_deg) ("biprism", z, def_x)
  ("kriv", z, f, coeffs) -- AberratedLensKrivanek, coeffs a dict of KRIVANEK_FIELDS

Pinning (tests/test_oracle_ray.py):
  * README golden: Ray(0.1, 0.2, 0.3, 0.4, 0, 0) through Lens(0.5, f=1) and
    Detector(1.0) -> (0.275, 0.4, 0.05, 0.0, z=1.0, pathlength=0.88875)
    (reference README.md:35-52), and its ABCD matrix (README.md:230-234);
  * closed forms of the reference tests (test_rays.py:26-39,
    test_component.py:186-259, :313-353);
  * pixel-transform known-answer tables (test_component.py:73-174) and the
    rotated scan-step vector (test_component.py:356-378);
  * binning: brute-force per-ray loop == vectorised counts; sum == in-range
    count; a planted off-by-one in the rounding fails the tables;
  * biprism: the reference's Jacobian tests (test_component.py:398-465: the
    deflection column on both sides of the apex, the deflection grown over
    the propagation, the lens + biprism system against the product of the
    analytic 5x5 matrices of tests/transfer_matrices.py), a hand-computed
    pathlength with the incoming slopes, and the identity with a Deflector
    of +-def_x away from the apex; planted mistakes (sign(-x), outgoing
    slopes in the pathlength) fail.
"""
from __future__ import annotations

import math

import numpy as np

FIELDS = ("x", "y", "dx", "dy", "z", "pathlength", "one")


def make_ray(x, y, dx, dy, z=0.0, pathlength=0.0, one=1.0):
    return dict(x=x, y=y, dx=dx, dy=dy, z=z, pathlength=pathlength, one=one)


def propagate(r, d):
    """FreeSpaceParaxial (reference propagator.py:51-72)."""
    return dict(r, x=r["x"] + r["dx"] * d, y=r["y"] + r["dy"] * d, z=r["z"] + d,
                pathlength=r["pathlength"] + d)


def lens(r, f):
    """Lens (reference components.py:161-174)."""
    x, y = r["x"], r["y"]
    return dict(r, dx=-x / f + r["dx"], dy=-y / f + r["dy"],
                pathlength=r["pathlength"] - (x * x + y * y) / (2 * f), one=r["one"] * 1.0)


def scanner(r, spx, spy, stx=0.0, sty=0.0):
    """Scanner (reference components.py:279-285)."""
    o = r["one"]
    return dict(r, x=r["x"] + spx * o, y=r["y"] + spy * o, dx=r["dx"] + stx * o, dy=r["dy"] + sty * o)


def descanner(r, spx, spy, stx, sty, e):
    """Descanner with the 12 DescanError coefficients in NamedTuple order (reference components.py:320-372)."""
    o = r["one"]
    return dict(
        r,
        x=r["x"] + (spx * e[0] + spy * e[1] + e[8] - spx) * o,
        y=r["y"] + (spx * e[2] + spy * e[3] + e[9] - spy) * o,
        dx=r["dx"] + (spx * e[4] + spy * e[5] + e[10] - stx) * o,
        dy=r["dy"] + (spx * e[6] + spy * e[7] + e[11] - sty) * o,
    )


def deflector(r, defx, defy):
    """Deflector; pathlength uses incoming slopes (reference components.py:476-482)."""
    x, y, dx, dy = r["x"], r["y"], r["dx"], r["dy"]
    o = r["one"]
    return dict(r, dx=dx + defx * o, dy=dy + defy * o, pathlength=r["pathlength"] + dx * x + dy * y)


def rotator(r, angle_deg):
    """Rotator (reference components.py:503-523); cos/sin of the angle in f64."""
    a = angle_deg * (math.pi / 180)
    c, s = math.cos(a), math.sin(a)
    return dict(r, x=r["x"] * c - r["y"] * s, y=r["x"] * s + r["y"] * c,
                dx=r["dx"] * c - r["dy"] * s, dy=r["dx"] * s + r["dy"] * c)


def biprism(r, def_x):
    """Biprism (reference components.py:553-559)."""
    x, y, dx, dy = r["x"], r["y"], r["dx"], r["dy"]
    return dict(r, dx=dx + def_x * r["one"] * np.sign(x), pathlength=r["pathlength"] + dx * x + dy * y)


KRIVANEK_FIELDS = ("C10", "C12", "phi12", "C21", "phi21", "C23", "phi23", "C30", "C32", "phi32", "C34", "phi34",
                   "C41", "phi41", "C43", "phi43", "C45", "phi45", "C50", "C52", "phi52", "C54", "phi54",
                   "C56", "phi56")


def _kc(c, m, ph, ph0):
    return np.cos(m * (ph - ph0))


def _ks(c, m, ph, ph0):
    return np.sin(m * (ph - ph0))


def krivanek_brackets(phi, c):
    """B2..B6 (reference aberrations.py:42-48); c: dict of KRIVANEK_FIELDS (missing = 0)."""
    g = lambda k: c.get(k, 0.0)  # noqa: E731
    B2 = g("C10") + g("C12") * _kc(c, 2, phi, g("phi12"))
    B3 = g("C21") * _kc(c, 1, phi, g("phi21")) + g("C23") * _kc(c, 3, phi, g("phi23"))
    B4 = g("C30") + g("C32") * _kc(c, 2, phi, g("phi32")) + g("C34") * _kc(c, 4, phi, g("phi34"))
    B5 = (g("C41") * _kc(c, 1, phi, g("phi41")) + g("C43") * _kc(c, 3, phi, g("phi43"))
          + g("C45") * _kc(c, 5, phi, g("phi45")))
    B6 = (g("C50") + g("C52") * _kc(c, 2, phi, g("phi52")) + g("C54") * _kc(c, 4, phi, g("phi54"))
          + g("C56") * _kc(c, 6, phi, g("phi56")))
    return B2, B3, B4, B5, B6


def krivanek_W(alpha, phi, c):
    """W_krivanek (reference aberrations.py:51-61)."""
    B2, B3, B4, B5, B6 = krivanek_brackets(phi, c)
    a = alpha
    a2 = a * a
    a3 = a2 * a
    a4 = a2 * a2
    a6 = a3 * a3
    return 0.5 * a2 * B2 + (a3 / 3.0) * B3 + 0.25 * a4 * B4 + 0.2 * a4 * a * B5 + (a6 / 6.0) * B6


def krivanek_grad(ax, ay, c):
    """grad_W_krivanek (reference aberrations.py:64-108): dW/dalpha and
    dW/dphi of the polar series, mapped to (ax, ay) with alpha -> 1e-30 at 0."""
    g = lambda k: c.get(k, 0.0)  # noqa: E731
    a = np.hypot(ax, ay)
    phi = np.arctan2(ay, ax)
    B2, B3, B4, B5, B6 = krivanek_brackets(phi, c)
    a2 = a * a
    a3 = a2 * a
    a4 = a2 * a2
    a5 = a4 * a
    a6 = a3 * a3
    dW_da = a * B2 + a2 * B3 + a3 * B4 + a4 * B5 + a5 * B6
    dW_dphi = (0.5 * a2) * (-2.0 * g("C12") * _ks(c, 2, phi, g("phi12")))
    dW_dphi = dW_dphi + (a3 / 3.0) * (-1.0 * g("C21") * _ks(c, 1, phi, g("phi21"))
                                      - 3.0 * g("C23") * _ks(c, 3, phi, g("phi23")))
    dW_dphi = dW_dphi + (0.25 * a4) * (-2.0 * g("C32") * _ks(c, 2, phi, g("phi32"))
                                       - 4.0 * g("C34") * _ks(c, 4, phi, g("phi34")))
    dW_dphi = dW_dphi + (0.2 * a4 * a) * (-1.0 * g("C41") * _ks(c, 1, phi, g("phi41"))
                                          - 3.0 * g("C43") * _ks(c, 3, phi, g("phi43"))
                                          - 5.0 * g("C45") * _ks(c, 5, phi, g("phi45")))
    dW_dphi = dW_dphi + (a6 / 6.0) * (-2.0 * g("C52") * _ks(c, 2, phi, g("phi52"))
                                      - 4.0 * g("C54") * _ks(c, 4, phi, g("phi54"))
                                      - 6.0 * g("C56") * _ks(c, 6, phi, g("phi56")))
    a_safe = np.where(a == 0, 1e-30, a)
    inv_a = 1.0 / a_safe
    inv_a2 = inv_a * inv_a
    dWx = dW_da * (ax * inv_a) + dW_dphi * (-ay * inv_a2)
    dWy = dW_da * (ay * inv_a) + dW_dphi * (ax * inv_a2)
    return dWx, dWy


def aberrated_lens(r, f, c):
    """AberratedLensKrivanek (reference components.py:192-214)."""
    x, y = r["x"], r["y"]
    idx = -x / f + r["dx"]
    idy = -y / f + r["dy"]
    alpha = np.hypot(idx, idy)
    phi = np.arctan2(idy, idx)
    dWx, dWy = krivanek_grad(idx, idy, c)
    return dict(r, dx=idx + (-dWx / f), dy=idy + (-dWy / f),
                pathlength=r["pathlength"] - (x * x + y * y) / (2 * f) + krivanek_W(alpha, phi, c) / f,
                one=r["one"] * 1.0)


def multi_cumsum_inplace(values, partitions, start):
    """Partitioned cumulative sum, restated loop for loop (reference utils.py:46-80)."""
    part_idx = 0
    current = partitions[part_idx]
    count = 0
    values[0] = start
    for v in range(1, values.size):
        if current == count:
            count = 0
            part_idx += 1
            current = partitions[part_idx]
            values[v] = start
        else:
            values[v] += values[v - 1]
            count += 1


def ring_params(num_points_approx, radius):
    """(radius, angle) of every point of concentric_rings (reference utils.py:117-175)."""
    num_rings = max(1, int(np.floor((-1 + np.sqrt(1 + 4 * num_points_approx / np.pi)) / 2)))
    kth = np.round(2 * np.pi * np.arange(1, num_rings + 1)).astype(int)
    num_rings = kth.size
    per_unit = num_points_approx / kth.sum()
    per_ring = np.round(kth * per_unit).astype(int)
    radii = np.linspace(0, radius, num_rings + 1, endpoint=True)[1:]
    with np.errstate(divide="ignore"):
        div_angle = 2 * np.pi / per_ring
    params = np.repeat(np.stack((radii, div_angle), axis=0), per_ring.tolist(), axis=-1)
    multi_cumsum_inplace(params[1, :], per_ring, 0.0)
    return params[0, :], params[1, :]


def concentric_rings(num_points_approx, radius):
    """(y, x) = (r sin a, r cos a) per point (reference utils.py:117-175)."""
    r, a = ring_params(num_points_approx, radius)
    return np.stack((r * np.sin(a), r * np.cos(a)), axis=-1)


def point_source_array(num, semi_conv, offset_xy=(0.0, 0.0)):
    """PointSource.generate_array(random=False) rows [x, y, dx, dy, 1] (reference source.py:104-135)."""
    dy, dx = concentric_rings(num, semi_conv).T
    r = np.zeros((dx.size, 5))
    r[:, 0] += offset_xy[0]
    r[:, 1] += offset_xy[1]
    r[:, 2] = dx
    r[:, 3] = dy
    r[:, 4] = 1.0
    return r


def parallel_beam_array(num, radius, offset_xy=(0.0, 0.0)):
    """ParallelBeam.generate_array(random=False) rows (reference source.py:156-188)."""
    y, x = concentric_rings(num, radius).T
    r = np.zeros((x.size, 5))
    r[:, 0] = x + offset_xy[0]
    r[:, 1] = y + offset_xy[1]
    r[:, 4] = 1.0
    return r


def fibonacci_spiral(nb_samples, radius, alpha=2):
    """Fibonacci-spiral disc points (reference utils.py:297-325)."""
    ga = np.pi * (3.0 - np.sqrt(5.0))
    boundary = np.round(alpha * np.sqrt(nb_samples))
    i = np.arange(nb_samples)
    rr = np.where(i > nb_samples - (boundary + 1), radius,
                  radius * np.sqrt((i + 0.5) / (nb_samples - 0.5 * (boundary + 1))))
    rr[0] = 0.0
    return rr * np.cos(i * ga), rr * np.sin(i * ga)


def thick_lens(r, z_po, z_pi, f):
    """ThickLens (reference components.py:410-451): the thin-lens update at its
    object plane, then the ray's z moves to the image plane."""
    r = lens(r, f)
    r["z"] = r["z"] - (z_po - z_pi)
    return r


def apply(r, comp):
    kind, *p = comp
    if kind in ("plane", "detector"):
        return r
    if kind == "lens":
        return lens(r, p[1])
    if kind == "scanner":
        return scanner(r, *p[1:])
    if kind == "descanner":
        return descanner(r, *p[1:])
    if kind == "deflector":
        return deflector(r, *p[1:])
    if kind == "rotator":
        return rotator(r, p[1])
    if kind == "kriv":
        return aberrated_lens(r, p[1], p[2])
    if kind == "biprism":
        return biprism(r, p[1])
    if kind == "thicklens":  # ("thicklens", z_po, z_pi, f)
        return thick_lens(r, p[0], p[1], p[2])
    raise ValueError(kind)


def run_to_end(r, model):
    """Step structure of reference run.py:75-82: propagate (even by 0) before every element."""
    for comp in model:
        r = propagate(r, comp[1] - r["z"])
        r = apply(r, comp)
    return r


def run_steps(r, model):
    """All intermediate states of run_iter (reference run.py:47-82)."""
    out = []
    for comp in model:
        r = propagate(r, comp[1] - r["z"])
        out.append(r)
        r = apply(r, comp)
        out.append(r)
    return out


# ---- pixel transforms (reference coordinate_transforms.py:7-140, grid.py:37-63) ----

def pixels_to_metres_matrix(centre, pixel_size, shape, flip_y=False, rotation=0.0):
    def rot(rad):
        return np.array([[math.cos(rad), math.sin(rad), 0.0], [-math.sin(rad), math.cos(rad), 0.0], [0, 0, 1.0]])

    def scale(p):
        return np.array([[p[0], 0.0, 0.0], [0.0, p[1], 0.0], [0, 0, 1.0]])

    def shift(c):
        return np.array([[1.0, 0.0, c[0]], [0.0, 1.0, c[1]], [0, 0, 1.0]])

    flip = np.array([[-1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    cond = flip if flip_y else np.eye(3)
    s = np.asarray(shape, dtype=np.float64)
    return cond @ flip @ rot(math.pi / 180 * rotation) @ shift(centre) @ scale(pixel_size) @ shift(-(s - 1) / 2.0)


def metres_to_pixels_matrix(centre, pixel_size, shape, flip_y=False, rotation=0.0):
    return np.linalg.inv(pixels_to_metres_matrix(centre, pixel_size, shape, flip_y, rotation))


def transform(y, x, T):
    """T @ [y, x, 1] in f64, summed in index order (reference coordinate_transforms.py:135-140)."""
    y = np.asarray(y, dtype=np.float64)
    x = np.asarray(x, dtype=np.float64)
    return (T[0, 0] * y + T[0, 1] * x) + T[0, 2], (T[1, 0] * y + T[1, 1] * x) + T[1, 2]


def round_to_i32(v):
    """jnp.round (half to even) then astype(int32) with XLA semantics (NaN -> 0, saturating)."""
    r = np.rint(np.asarray(v, dtype=np.float64))
    r = np.where(np.isnan(r), 0.0, r)
    return np.clip(r, -2147483648.0, 2147483647.0).astype(np.int32)


def metres_to_pixels(x, y, Tinv):
    py, px = transform(y, x, Tinv)
    return round_to_i32(py), round_to_i32(px)


def into_image(py, px, shape, acc=None):
    """Counting scatter of reference utils.py:83-114 (bounds-checked, out-of-range dropped)."""
    h, w = int(shape[0]), int(shape[1])
    acc = np.zeros((h, w), dtype=np.int64) if acc is None else acc
    py = np.ravel(py).astype(np.int64)
    px = np.ravel(px).astype(np.int64)
    ok = (py >= 0) & (py < h) & (px >= 0) & (px < w)
    np.add.at(acc, (py[ok], px[ok]), 1)
    return acc


def into_image_loop(py, px, shape):
    """Serial per-ray loop, the literal numba algorithm (for brute-force checks)."""
    h, w = int(shape[0]), int(shape[1])
    acc = np.zeros((h, w), dtype=np.int64)
    for a, b in zip(np.ravel(py), np.ravel(px)):
        if 0 <= a < h and 0 <= b < w:
            acc[a, b] += 1
    return acc


# ---- the plain run_to_end workload (lens + detector, uniform random rays) ----

def plain_workload(n, seed=0, dtype=np.float32):
    """Seeded rays: draw order x, y, dx, dy from default_rng(seed), in f64 then cast."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.1, 0.1, n).astype(dtype)
    y = rng.uniform(-0.1, 0.1, n).astype(dtype)
    dx = rng.uniform(-0.05, 0.05, n).astype(dtype)
    dy = rng.uniform(-0.05, 0.05, n).astype(dtype)
    return x, y, dx, dy


PLAIN_MODEL = [("lens", 0.5, 1.0), ("detector", 1.0)]
PLAIN_DET = dict(centre=(0.0, 0.0), pixel_size=(0.01, 0.01), shape=(128, 128))


def plain_run(x, y, dx, dy):
    """Plain workload through the oracle: (ray dict, int64 histogram)."""
    r = run_to_end(make_ray(x, y, dx, dy), PLAIN_MODEL)
    Tinv = metres_to_pixels_matrix(**PLAIN_DET)
    py, px = metres_to_pixels(r["x"], r["y"], Tinv)
    return r, into_image(py, px, PLAIN_DET["shape"])
