"""AberratedLensKrivanek on the host (no GPU): the polynomial form of the
Krivanek series against the oracle's trigonometric restatement of the
reference (aberrations.py), the scalar-ray path, the compiled ray program
(interpreted here one IEEE operation per instruction) and the 5x5 matrices.

Tolerances: the polynomial and trigonometric forms round differently, so
values are compared to 1e-12 (f64) or 2e-6 (f32) of the size of the terms
they sum, not bit for bit."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle import ray_oracle as O
from temgymcore_amd import _program
from temgymcore_amd.aberrations import KrivanekCoeffs, W_krivanek, grad_W_krivanek, krivanek_poly
from temgymcore_amd.components import AberratedLensKrivanek, Detector, Lens
from temgymcore_amd.propagator import FreeSpaceParaxial
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import _trace_model, run_to_end
from temgymcore_amd.transfer import ray_matrix
from test_host_compiler import interpret


def _coeffs(rng, scale=1.0):
    d = {}
    for f in dataclasses.fields(KrivanekCoeffs):
        d[f.name] = rng.uniform(-np.pi, np.pi) if f.name.startswith("phi") else scale * rng.uniform(-1, 1)
    return d


def _term_scale(ax, ay, c):
    """sum over terms of |C| a^(n+1) and of its slope derivative |C| a^n (n+1)."""
    a = np.hypot(ax, ay)
    orders = {"C10": 1, "C12": 1, "C21": 2, "C23": 2, "C30": 3, "C32": 3, "C34": 3, "C41": 4, "C43": 4,
              "C45": 4, "C50": 5, "C52": 5, "C54": 5, "C56": 5}
    w = sum(abs(c.get(k, 0.0)) * a ** (n + 1) for k, n in orders.items())
    g = sum(abs(c.get(k, 0.0)) * (n + 1) * a ** n for k, n in orders.items())
    return w, g


def test_polynomial_form_matches_trigonometric_series():
    rng = np.random.default_rng(51)
    c = _coeffs(rng)
    ax, ay = rng.uniform(-0.7, 0.7, 400), rng.uniform(-0.7, 0.7, 400)
    ax[0] = ay[0] = 0.0  # alpha = 0: the reference's 1e-30 guard
    W, gx, gy = krivanek_poly(ax, ay, c)
    sw, sg = _term_scale(ax, ay, c)
    W_ref = O.krivanek_W(np.hypot(ax, ay), np.arctan2(ay, ax), c)
    gx_ref, gy_ref = O.krivanek_grad(ax, ay, c)
    assert np.all(np.abs(W - W_ref) <= 1e-13 * sw + 1e-300)
    assert np.all(np.abs(gx - gx_ref) <= 1e-13 * sg + 1e-300)
    assert np.all(np.abs(gy - gy_ref) <= 1e-13 * sg + 1e-300)
    assert gx[0] == 0.0 and gy[0] == 0.0
    # the reference-named wrappers
    np.testing.assert_allclose(W_krivanek(np.hypot(ax, ay), np.arctan2(ay, ax), KrivanekCoeffs(**c)), W_ref,
                               rtol=1e-12, atol=1e-14)
    g2 = grad_W_krivanek(ax, ay, c)
    np.testing.assert_array_equal(g2[0], gx)


@pytest.mark.parametrize("only", ["C10", "C12", "C21", "C23", "C30", "C32", "C34", "C41", "C43", "C45",
                                  "C50", "C52", "C54", "C56"])
def test_each_term_alone(only):
    """Each (order, fold) term separately, so one wrong power or fold cannot
    hide behind the others."""
    rng = np.random.default_rng(52)
    c = {only: 0.7}
    ph = "phi" + only[1:]
    if ph in {f.name for f in dataclasses.fields(KrivanekCoeffs)}:
        c[ph] = 0.37
    ax, ay = rng.uniform(-0.7, 0.7, 100), rng.uniform(-0.7, 0.7, 100)
    W, gx, gy = krivanek_poly(ax, ay, c)
    sw, sg = _term_scale(ax, ay, c)
    W_ref = O.krivanek_W(np.hypot(ax, ay), np.arctan2(ay, ax), c)
    gx_ref, gy_ref = O.krivanek_grad(ax, ay, c)
    assert np.all(np.abs(W - W_ref) <= 1e-13 * sw)
    assert np.all(np.abs(gx - gx_ref) <= 1e-13 * sg)
    assert np.all(np.abs(gy - gy_ref) <= 1e-13 * sg)


def _model(c, f=2.5e-4):
    return [AberratedLensKrivanek(z=f, focal_length=f, coeffs=c),
            Detector(z=2 * f, pixel_size=(1e-6, 1e-6), shape=(8, 8))]


def test_scalar_ray_vs_oracle():
    rng = np.random.default_rng(53)
    c = _coeffs(rng, scale=1e4)
    for _ in range(5):
        x, y, dx, dy = rng.uniform(-2e-8, 2e-8, 2).tolist() + rng.uniform(-1e-5, 1e-5, 2).tolist()
        out = run_to_end(Ray(x, y, dx, dy, 0.0, 0.0), _model(c))
        ref = O.run_to_end(O.make_ray(x, y, dx, dy), [("kriv", 2.5e-4, 2.5e-4, c), ("detector", 5e-4)])
        for k in ("x", "y", "dx", "dy", "pathlength"):
            assert getattr(out, k) == pytest.approx(float(ref[k]), rel=1e-11, abs=1e-24), k


def _run_host_split(ray, model):
    """Trace, compile (splitting where needed) and interpret every program."""
    tr = _program.Trace(torch.device("cpu"))
    out = _trace_model(tr.symbolize(ray), model, FreeSpaceParaxial())
    names = [n for n in ("x", "y", "dx", "dy", "pathlength")
             if isinstance(getattr(out, n), _program.Sym) and getattr(out, n).vid not in tr.value_slot]
    vids = [getattr(out, n).vid for n in names]
    prefixes, last, prefix_comps = _program._compile_segments(tr, vids, [], _program._shape_of(tr, vids[0]))
    got = {}
    for comp, outs in prefix_comps:
        vals = interpret(tr, comp)
        for v, t in outs.items():
            t.copy_(torch.from_numpy(vals[v]))
        got.update(vals)
    got.update(interpret(tr, last))
    return {n: got[v] for n, v in zip(names, vids)}, len(prefixes) + 1


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_compiled_program_vs_oracle(dtype):
    """Every Krivanek term at once: the traced program is long enough to be
    split into several programs."""
    rng = np.random.default_rng(54)
    n = 3000
    c = _coeffs(rng, scale=1e4)
    x, y = (rng.uniform(-2e-8, 2e-8, n).astype(dtype) for _ in range(2))
    dx, dy = (rng.uniform(-1e-5, 1e-5, n).astype(dtype) for _ in range(2))
    ray = Ray(x=torch.from_numpy(x), y=torch.from_numpy(y), dx=torch.from_numpy(dx), dy=torch.from_numpy(dy),
              z=0.0, pathlength=0.0)
    res, n_prog = _run_host_split(ray, _model(c))
    assert n_prog >= 2
    ref = O.run_to_end(O.make_ray(x.astype(np.float64), y.astype(np.float64), dx.astype(np.float64),
                                  dy.astype(np.float64)), [("kriv", 2.5e-4, 2.5e-4, c), ("detector", 5e-4)])
    tol = 1e-12 if dtype == np.float64 else 2e-6
    f = 2.5e-4
    idx, idy = -x.astype(np.float64) / f + dx, -y.astype(np.float64) / f + dy
    _, sg = _term_scale(idx, idy, c)
    for k in ("dx", "dy"):
        scale = np.abs(idx) + np.abs(idy) + sg / f
        assert res[k].dtype == dtype
        assert np.all(np.abs(res[k] - ref[k]) <= tol * scale), k
    for k in ("x", "y"):
        scale = np.abs(x) + f * (np.abs(idx) + np.abs(idy) + sg / f)
        assert np.all(np.abs(res[k] - ref[k]) <= tol * scale), k


def test_notebook_golden_through_compiled_program():
    """examples/aberrated_probe.ipynb cell 9 through the compiled f64 program."""
    import json
    import os

    with open(os.path.join(os.path.dirname(__file__), "golden", "aberrated_probe_rays.json")) as fh:
        g = json.load(fh)
    st = g["setup"]
    x, y = O.fibonacci_spiral(st["n_beams"], st["radius"], alpha=st["alpha"])
    n, F1 = len(x), st["F1"]
    z = torch.zeros(n, dtype=torch.float64)
    ray = Ray(x=torch.from_numpy(x), y=torch.from_numpy(y), dx=z, dy=z, z=z, pathlength=z)
    model = [AberratedLensKrivanek(z=F1, focal_length=F1, coeffs=KrivanekCoeffs(**st["coeffs"])),
             Detector(z=2 * F1, pixel_size=(1e-6, 1e-6), shape=(8, 8))]
    res, _ = _run_host_split(ray, model)
    idx = np.array(g["index"])
    for k in ("x", "y", "dx", "dy"):
        np.testing.assert_allclose(res[k][idx], g[k], rtol=6e-9, atol=0, err_msg=k)


def test_ray_matrix_vs_oracle_differences():
    """The 5x5 matrix of an aberrated lens depends on the ray; the dual-number
    Jacobian equals central differences of the oracle."""
    rng = np.random.default_rng(55)
    c = _coeffs(rng, scale=1e4)
    model = [Lens(z=1e-4, focal_length=3e-4)] + _model(c)
    model_t = [("lens", 1e-4, 3e-4), ("kriv", 2.5e-4, 2.5e-4, c), ("detector", 5e-4)]
    base = np.array([1.3e-8, -0.7e-8, 2e-6, -3e-6, 1.0])
    J = np.asarray(ray_matrix(model, Ray(*base[:4].tolist(), 0.0, 0.0)))
    ref = np.zeros((5, 5))
    for j in range(5):
        h = 1e-4 * abs(base[j])

        def run(v):
            r = O.run_to_end(O.make_ray(v[0], v[1], v[2], v[3], 0.0, 0.0, v[4]), model_t)
            return np.array([r["x"], r["y"], r["dx"], r["dy"], r["one"]], dtype=np.float64)

        def cd(hh):
            a, b = base.copy(), base.copy()
            a[j] += hh
            b[j] -= hh
            return (run(a) - run(b)) / (2 * hh)

        ref[:, j] = (4 * cd(h / 2) - cd(h)) / 3
    scale = np.abs(ref).max(axis=1, keepdims=True)
    assert np.all(np.abs(J - ref) <= 1e-7 * scale + 1e-300)
    assert abs(J[2, 0] - (-1 / 3e-4)) > 1e-3 / 3e-4  # the aberration changes the matrix


def test_coefficients_from_dict_and_params_path():
    lens = AberratedLensKrivanek(z=0.1, focal_length=0.2, coeffs={"C30": 1.0})
    assert isinstance(lens.coeffs, KrivanekCoeffs) and lens.coeffs.C30 == 1.0
    path = lens.params.coeffs.C30
    root, *keys = path._path()
    assert root is lens and keys == ["coeffs", "C30"]
    with pytest.raises(TypeError):
        AberratedLensKrivanek(z=0.1, focal_length=0.2, coeffs=[1.0])


@pytest.mark.parametrize("name", ["C30", "C10", "phi12", "C56"])
def test_run_with_grads_coefficients_scalar_ray(name):
    """d(output)/d(coefficient) for scalar rays, including coefficients whose
    value is 0 (their terms must still carry a tangent)."""
    from temgymcore_amd.run import run_with_grads

    f = 2.5e-4
    c = {"C30": 1e6, "C12": 1e-3, "phi12": 0.2}
    lens = AberratedLensKrivanek(z=f, focal_length=f, coeffs=c)
    model = [lens, Detector(z=2 * f, pixel_size=(1e-6, 1e-6), shape=(8, 8))]
    r0 = (1.1e-8, -2.3e-8, 1.7e-6, -0.4e-6)
    _, grads = run_with_grads(Ray(*r0, 0.0, 0.0), model, [getattr(lens.params.coeffs, name)])
    g = grads[(lens, "coeffs", name)]
    v0 = c.get(name, 0.0)
    h = 1e-3 * abs(v0) if v0 else {"C10": 1e-6, "C56": 1e6}.get(name, 1e-3)

    def run(v):
        return O.run_to_end(O.make_ray(*r0, 0.0, 0.0), [("kriv", f, f, dict(c, **{name: v})), ("detector", 2 * f)])

    def cd(hh):
        a, b = run(v0 + hh), run(v0 - hh)
        return {k: (a[k] - b[k]) / (2 * hh) for k in ("x", "y", "dx", "dy", "pathlength")}

    d1, d2 = cd(h), cd(h / 2)
    refs = {k: (4 * d2[k] - d1[k]) / 3 for k in d1}
    scale = max(abs(refs[k]) for k in ("dx", "dy"))
    assert scale > 0
    for k in ("dx", "dy"):
        assert getattr(g, k) == pytest.approx(refs[k], rel=1e-6, abs=1e-6 * scale), k


def test_fibonacci_spiral_matches_oracle():
    from temgymcore_amd.source import fibonacci_spiral

    for n, alpha in ((100000, 0), (5000, 2)):
        a = fibonacci_spiral(n, 0.2e-9, alpha=alpha)
        b = O.fibonacci_spiral(n, 0.2e-9, alpha=alpha)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
