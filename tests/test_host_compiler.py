"""Host-side checks (no GPU): the component tracer / register allocator, the
library's exported symbols, and the scalar-ray API against the oracle.

The compiled register programs are executed here by a small numpy
interpreter that applies one IEEE operation per instruction, exactly as the
kernel does; its results must be bit-identical to the oracle's f32 run."""
import ctypes
import dataclasses
import os
import re

import numpy as np
import pytest
import torch

from oracle import ray_oracle as O
from temgymcore_amd import _lib, _program
from temgymcore_amd.components import (
    Biprism, Deflector, DescanError, Descanner, Detector, Lens, Plane, Rotator, Scanner,
)
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import _split, _trace_model, run_iter, run_to_end, solve_model
from temgymcore_amd.propagator import FreeSpaceParaxial
from temgymcore_amd.transfer import custom_jacobian_matrix, jacobian, ray_matrix

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INV_OPC = {v: k for k, v in _lib.OPCODES.items()}


def interpret(trace, comp):
    """numpy execution of a Compiled program (test infrastructure)."""
    dt = np.float32 if comp.dtype == torch.float32 else np.float64
    shape = tuple(comp.shape)
    regs = [np.zeros(shape, dt) for _ in range(_lib.TG_MAX_REG)]
    for s, r in comp.slots:
        regs[r] = np.broadcast_to(trace.slot_tensors[s].numpy().astype(dt), shape).copy()
    for op, d, a, b, imm in comp.instrs:
        narrow = bool(op & _lib.TG_OP_F32)
        op &= 0x7F
        im = np.float64(np.float32(imm)) if narrow else dt(imm)
        A, B = regs[a], regs[b]
        name = INV_OPC[op]
        with np.errstate(all="ignore"):
            val = {
                "add": lambda: A + B, "sub": lambda: A - B, "mul": lambda: A * B, "div": lambda: A / B,
                "addi": lambda: A + im, "subi": lambda: A - im, "rsubi": lambda: im - A,
                "muli": lambda: A * im, "divi": lambda: A / im, "rdivi": lambda: im / A,
                "neg": lambda: -A, "mov": lambda: A.copy(), "set": lambda: np.full(shape, im, dt),
                "sign": lambda: np.sign(A),
            }[name]()
        regs[d] = val.astype(np.float32).astype(dt) if narrow else val.astype(dt)
    return {v: regs[r].astype(np.float32 if trace.vdtype[v] == torch.float32 else dt) for v, r in comp.outputs}


def run_host(ray, model):
    """Trace + compile `model` on CPU tensors and interpret: dict of outputs."""
    tr = _program.Trace(torch.device("cpu"))
    out = _trace_model(tr.symbolize(ray), model, FreeSpaceParaxial())
    res = {}
    groups = {}
    for name in ("x", "y", "dx", "dy", "z", "pathlength", "_one"):
        v = getattr(out, name)
        if isinstance(v, _program.Sym):
            if v.vid in tr.value_slot:
                res[name] = tr.slot_tensors[tr.value_slot[v.vid]].numpy()
            else:
                groups.setdefault(_program._shape_of(tr, v.vid), []).append((name, v.vid))
        else:
            res[name] = v
    for shape, items in groups.items():
        comp = _program.compile_group(tr, [vid for _, vid in items], [], shape)
        vals = interpret(tr, comp)
        for name, vid in items:
            res[name] = vals[vid]
    return res, tr


def test_plain_model_bit_exact_vs_oracle():
    x, y, dx, dy = O.plain_workload(50000, seed=0)
    ref, _ = O.plain_run(x, y, dx, dy)
    ray = Ray(x=torch.from_numpy(x), y=torch.from_numpy(y), dx=torch.from_numpy(dx), dy=torch.from_numpy(dy),
              z=0.0, pathlength=0.0)
    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]
    res, _ = run_host(ray, model)
    for k in ("x", "y", "dx", "dy", "pathlength"):
        assert res[k].dtype == np.float32
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)
    assert res["z"] == ref["z"] == 1.0


def _rich_model(err):
    return [
        Scanner(z=0.1, scan_pos_x=0.003, scan_pos_y=-0.002, scan_tilt_x=1e-3, scan_tilt_y=-2e-3),
        Deflector(z=0.2, def_x=0.01, def_y=-0.02),
        Lens(z=0.35, focal_length=0.7),
        Rotator(z=0.4, angle=12.0),
        Descanner(z=0.5, scan_pos_x=0.003, scan_pos_y=-0.002, scan_tilt_x=1e-3, scan_tilt_y=-2e-3,
                  descan_error=DescanError(*err)),
        Biprism(z=0.55, def_x=1e-3),
        Plane(z=0.6),
        Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(64, 64)),
    ]


def _rich_oracle(err):
    return [
        ("scanner", 0.1, 0.003, -0.002, 1e-3, -2e-3), ("deflector", 0.2, 0.01, -0.02),
        ("lens", 0.35, 0.7), ("rotator", 0.4, 12.0),
        ("descanner", 0.5, 0.003, -0.002, 1e-3, -2e-3, err), ("biprism", 0.55, 1e-3),
        ("plane", 0.6), ("detector", 1.0),
    ]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rich_model_bit_exact_vs_oracle(dtype):
    rng = np.random.default_rng(7)
    n = 4097
    x, y = (rng.uniform(-0.05, 0.05, n).astype(dtype) for _ in range(2))
    dx, dy = (rng.uniform(-0.02, 0.02, n).astype(dtype) for _ in range(2))
    pl = rng.uniform(0, 1e-3, n).astype(dtype)
    err = tuple(float(v) for v in rng.normal(0, 1e-3, 12))
    ray = Ray(x=torch.from_numpy(x), y=torch.from_numpy(y), dx=torch.from_numpy(dx), dy=torch.from_numpy(dy),
              z=0.05, pathlength=torch.from_numpy(pl))
    res, _ = run_host(ray, _rich_model(err))
    ref = O.run_to_end(O.make_ray(x, y, dx, dy, 0.05, pl), _rich_oracle(err))
    for k in ("x", "y", "dx", "dy", "pathlength"):
        assert res[k].dtype == dtype
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


def test_point_source_scalar_positions_weak_typing():
    # x, y uniform (Python floats), slopes per ray: the first propagation must
    # round the scalar x once to f32 and add fl(dx*d) -- as the oracle does.
    rng = np.random.default_rng(1)
    dx = rng.uniform(-1e-3, 1e-3, 999).astype(np.float32)
    dy = rng.uniform(-1e-3, 1e-3, 999).astype(np.float32)
    ray = Ray(x=1e-9, y=-3e-9, dx=torch.from_numpy(dx), dy=torch.from_numpy(dy), z=0.0, pathlength=0.0)
    model = [Lens(z=2e-7, focal_length=1e-3), Detector(z=0.5, pixel_size=(1e-5, 1e-5), shape=(32, 32))]
    res, _ = run_host(ray, model)
    ref = O.run_to_end(O.make_ray(1e-9, -3e-9, dx, dy), [("lens", 2e-7, 1e-3), ("detector", 0.5)])
    for k in ("x", "y", "dx", "dy", "pathlength"):
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


def test_broadcast_scan_positions():
    # per-scan-position parameters [S,1] against per-ray slopes [1,J]
    S, J = 5, 7
    rng = np.random.default_rng(2)
    spx = rng.uniform(-1e-9, 1e-9, (S, 1))
    dx = rng.uniform(-1e-3, 1e-3, (1, J)).astype(np.float32)
    ray = Ray(x=0.0, y=0.0, dx=torch.from_numpy(dx), dy=torch.from_numpy(dx.copy()), z=0.0, pathlength=0.0)
    model = [Scanner(z=0.0, scan_pos_x=torch.from_numpy(spx), scan_pos_y=0.0), Plane(z=1e-7)]
    res, _ = run_host(ray, model)
    assert res["x"].shape == (S, J) and res["x"].dtype == np.float64  # f64 params promote
    ref = O.run_to_end(O.make_ray(0.0, 0.0, dx, dx.copy()), [("scanner", 0.0, spx, 0.0, 0.0, 0.0), ("plane", 1e-7)])
    np.testing.assert_array_equal(res["x"], ref["x"])
    assert res["dx"].shape == (1, J)  # unchanged field keeps its own shape


def test_register_pressure_and_dce():
    x = torch.zeros(10, dtype=torch.float32)
    ray = Ray(x=x, y=x.clone(), dx=x.clone(), dy=x.clone(), z=0.0, pathlength=0.0)
    model = [Lens(z=0.1 * (i + 1), focal_length=1.0 + i) for i in range(6)]
    tr = _program.Trace(torch.device("cpu"))
    out = _trace_model(tr.symbolize(ray), model, FreeSpaceParaxial())
    comp = _program.compile_group(tr, [out.x.vid], [], (10,))
    full = _program.compile_group(tr, [out.x.vid, out.pathlength.vid], [], (10,))
    assert len(comp.instrs) < len(full.instrs)  # pathlength work eliminated
    regs = {r for ins in full.instrs for r in ins[1:4]}
    assert max(regs) < _lib.TG_MAX_REG


def test_control_flow_on_per_ray_value_raises():
    class Bad(Plane):
        def __call__(self, ray):
            if ray.x > 0:
                return ray
            return ray

    tr = _program.Trace(torch.device("cpu"))
    ray = tr.symbolize(Ray(x=torch.zeros(3), y=0.0, dx=0.0, dy=0.0, z=0.0, pathlength=0.0))
    with pytest.raises(TypeError):
        Bad(z=0.0)(ray)


def test_scalar_api_matches_oracle_and_readme():
    r = Ray(x=0.1, y=0.2, dx=0.3, dy=0.4, z=0.0, pathlength=0.0)
    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]
    out = run_to_end(r, model)
    ref = O.run_to_end(O.make_ray(0.1, 0.2, 0.3, 0.4), [("lens", 0.5, 1.0), ("detector", 1.0)])
    assert (out.x, out.y, out.dx, out.dy, out.z, out.pathlength) == (
        ref["x"], ref["y"], ref["dx"], ref["dy"], ref["z"], ref["pathlength"])
    np.testing.assert_allclose(ray_matrix(model, r), [
        [0.5, 0, 0.75, 0, 0], [0, 0.5, 0, 0.75, 0], [-1, 0, 0.5, 0, 0], [0, -1, 0, 0.5, 0], [0, 0, 0, 0, 1]])
    # per-step matrices (reference README.md:245-267)
    steps = solve_model(r, model)
    assert steps.shape == (4, 5, 5)
    np.testing.assert_allclose(steps[0][0, 2], 0.5)
    np.testing.assert_allclose(steps[1][2, 0], -1.0)


def test_run_iter_step_structure():
    # reference test_run.py:10-47: a propagator step before every element, even at d = 0
    from temgymcore_amd.propagator import Propagator
    from temgymcore_amd.source import PointSource

    comps = (PointSource(z=0.0, semi_conv=0.023), Scanner(z=1.2, scan_pos_x=0.0, scan_pos_y=0.0), Plane(z=1.2),
             Descanner(z=1.2, scan_pos_x=0.0, scan_pos_y=0.0), Plane(z=3.1))
    ray = Ray(x=0.12, y=0.23, dx=0.34, dy=0.45, z=3.14, pathlength=0.34)
    res = list(run_iter(ray, comps))
    assert len(res) == 2 * len(comps)
    prev = ray
    for i, c in enumerate(comps):
        prop, pr = res[2 * i]
        comp, cr = res[2 * i + 1]
        assert isinstance(prop, Propagator) and isinstance(prop.propagator, FreeSpaceParaxial)
        assert comp is c
        assert prop.distance == c.z - prev.z
        assert pr.z == pytest.approx(c.z) and cr.z == pytest.approx(c.z)
        assert (pr.dx, pr.dy) == (prev.dx, prev.dy)
        assert pr.x == pytest.approx(prev.x + prev.dx * prop.distance)
        assert pr.y == pytest.approx(prev.y + prev.dy * prop.distance)
        prev = cr


def test_run_iter_without_propagation_composes_the_elements():
    # reference test_run.py:50-92: every element at one z, so each step is the
    # element alone and the result is the plain composition
    from temgymcore_amd.source import PointSource

    z = 1.2
    comps = (PointSource(z=z, semi_conv=0.023), Scanner(z=z, scan_pos_x=23.0, scan_pos_y=42.0), Plane(z=z),
             Descanner(z=z, scan_pos_x=13.0, scan_pos_y=11.0), Plane(z=z))
    ray = Ray(x=0.12, y=0.23, dx=0.34, dy=0.45, z=z, pathlength=0.34)
    res = list(run_iter(ray, comps))
    ref = ray
    for c in comps:
        ref = c(ref)
    for i, c in enumerate(comps):
        assert res[2 * i][0].distance == 0.0
        assert res[2 * i][1].z == pytest.approx(z) and res[2 * i + 1][1].z == pytest.approx(z)
    final = res[-1][1]
    for k in ("x", "y", "dx", "dy", "_one", "z", "pathlength"):
        assert getattr(final, k) == pytest.approx(getattr(ref, k), rel=1e-15, abs=1e-15), k
    assert (final.x, final.y) != (ray.x, ray.y)  # the scanner and descanner moved it


def test_descanner_jacobian_dual():
    e = np.random.default_rng(3).random(12)
    d = Descanner(z=0.0, scan_pos_x=1.5, scan_pos_y=-2.0, descan_error=DescanError(*e))
    J = custom_jacobian_matrix(jacobian(d)(Ray(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)))
    T = np.eye(5)
    T[0, 4] = 1.5 * e[0] - 2.0 * e[1] + e[8] - 1.5
    T[1, 4] = 1.5 * e[2] - 2.0 * e[3] + e[9] + 2.0
    T[2, 4] = 1.5 * e[4] - 2.0 * e[5] + e[10]
    T[3, 4] = 1.5 * e[6] - 2.0 * e[7] + e[11]
    np.testing.assert_allclose(J, T, atol=1e-12)


def test_scan_grid_tables_host():
    from temgymcore_amd.components import ScanGrid
    for xy, rot, expected in [((-0.5, 0.5), 90.0, (10, 0)), ((0.5, -0.5), 0.0, (10, 10))]:
        g = ScanGrid(z=0.0, rotation=rot, pixel_size=(0.1, 0.1), shape=(11, 11))
        py, px = g.metres_to_pixels(xy)
        assert (int(py), int(px)) == expected


# ---- C ABI library ---------------------------------------------------------

def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^[A-Za-z_][\w \*]*?\b(tg_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    from temgymcore_amd import gauss_fit, gaussian, stem  # noqa: F401  (register their signatures)

    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared_symbols()
    assert "tg_ray_program_run" in names
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_lib._SIGNATURES) >= names


def test_struct_layout_matches_header(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's layout."""
    import shutil
    import subprocess

    from temgymcore_amd import gauss_fit, gaussian, stem

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {
        "tg_instr": _lib.Instr, "tg_slot": _lib.Slot, "tg_program": _lib.Program, "tg_bin_spec": _lib.BinSpec,
        "tg_stem_geometry": stem.StemGeometry, "tg_beamlet_coef": gaussian.BeamletCoef,
        "tg_fit_instance": gauss_fit.FitInstance, "tg_fit_plan": gauss_fit.FitPlan,
    }
    lines, expect = [], []
    for cname, py in mirrors.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(ctypes.sizeof(py))
        for fname, *_ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(py, fname).offset)
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "temgym_hip.h"\nint main(){'
                   + "".join(lines) + "}")
    exe = tmp_path / "sz"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == expect


def _run_host_split(ray, model):
    """Like run_host, through the splitting compiler: every prefix program is
    interpreted into its buffers before the next one runs."""
    tr = _program.Trace(torch.device("cpu"))
    out = _trace_model(tr.symbolize(ray), model, FreeSpaceParaxial())
    names = [n for n in ("x", "y", "dx", "dy", "pathlength") if isinstance(getattr(out, n), _program.Sym)]
    vids = [getattr(out, n).vid for n in names]
    shape = _program._shape_of(tr, vids[0])
    prefixes, last, prefix_comps = _program._compile_segments(tr, vids, [], shape)
    for comp, outs in prefix_comps:
        vals = interpret(tr, comp)
        for v, t in outs.items():
            t.copy_(torch.from_numpy(vals[v]))
    vals = interpret(tr, last)
    return {n: vals[v] for n, v in zip(names, vids)}, len(prefixes)


def test_long_model_is_split_into_programs_bit_exact():
    """A model longer than one program (96 instructions) runs as several
    programs that hand values over through buffers, bit-exact to the oracle."""
    x, y, dx, dy = O.plain_workload(4000, seed=5)
    ray = Ray(x=torch.from_numpy(x), y=torch.from_numpy(y), dx=torch.from_numpy(dx), dy=torch.from_numpy(dy),
              z=0.0, pathlength=0.0)
    model, omodel = [], []
    for i in range(14):
        z = 0.05 * (i + 1)
        if i % 2:
            model.append(Lens(z=z, focal_length=0.7 + 0.01 * i))
            omodel.append(("lens", z, 0.7 + 0.01 * i))
        else:
            model.append(Deflector(z=z, def_x=1e-3 * i, def_y=-2e-4))
            omodel.append(("deflector", z, 1e-3 * i, -2e-4))
    model.append(Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8)))
    omodel.append(("detector", 1.0))
    tr = _program.Trace(torch.device("cpu"))
    out = _trace_model(tr.symbolize(ray), model, FreeSpaceParaxial())
    with pytest.raises((_program._TooLong, _program._TooManyRegs)):
        _program.compile_group(tr, [out.x.vid, out.y.vid, out.dx.vid, out.dy.vid, out.pathlength.vid], [], (4000,))
    res, n_prefix = _run_host_split(ray, model)
    assert n_prefix >= 1
    ref = O.run_to_end(O.make_ray(x, y, dx, dy), omodel)
    for k in ("x", "y", "dx", "dy", "pathlength"):
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


def test_ray_matrix_through_biprism_vs_oracle_differences():
    """The 5x5 ABCD of a model with a Biprism depends on the side of the
    biprism the central ray passes (reference components.py:553-559)."""
    from temgymcore_amd.components import Biprism

    model = [Lens(z=2e-4, focal_length=5e-4), Biprism(z=4e-4, def_x=1e-5),
             Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(8, 8))]
    model_t = [("lens", 2e-4, 5e-4), ("biprism", 4e-4, 1e-5), ("detector", 1e-3)]
    for x0 in (1e-7, -1e-7):
        J = np.asarray(ray_matrix(model, Ray(x0, 2e-8, 1e-5, 0.0, 0.0, 0.0)))
        base = np.array([x0, 2e-8, 1e-5, 0.0, 1.0])
        ref = np.zeros((5, 5))
        for j in range(5):
            outs = []
            for s in (1, -1):
                v = base.copy()
                v[j] += s * 1e-10
                r = O.run_to_end(O.make_ray(v[0], v[1], v[2], v[3], 0.0, 0.0, v[4]), model_t)
                outs.append(np.array([r["x"], r["y"], r["dx"], r["dy"], r["one"]]))
            ref[:, j] = (outs[0] - outs[1]) / 2e-10
        np.testing.assert_allclose(J, ref, rtol=1e-6, atol=1e-9)
        assert np.sign(J[2, 4]) == np.sign(x0)


def test_traced_jacobian_with_many_outputs_is_split_by_outputs(monkeypatch):
    """A forward-mode trace of a Biprism model stores 11-12 values (outputs
    and their tangents); with a program limit of 10 outputs the extra outputs
    get programs of their own.  Interpreted, every column equals the host
    matrix's."""
    monkeypatch.setattr(_lib, "TG_MAX_SLOT", 10)
    model = [Lens(z=2e-4, focal_length=5e-4), Biprism(z=4e-4, def_x=1e-5),
             Deflector(z=6e-4, def_x=1e-6, def_y=-2e-6), Detector(z=1e-3, pixel_size=(3e-8, 3e-8), shape=(8, 8))]
    rng = np.random.default_rng(32)
    n = 6
    cols = {k: torch.from_numpy(rng.uniform(-2e-7, 2e-7, n)) for k in ("x", "y")}
    cols.update({k: torch.from_numpy(rng.uniform(-1e-5, 1e-5, n)) for k in ("dx", "dy")})
    ray = Ray(x=cols["x"], y=cols["y"], dx=cols["dx"], dy=cols["dy"], z=torch.zeros(n, dtype=torch.float64),
              pathlength=torch.zeros(n, dtype=torch.float64), _one=torch.ones(n, dtype=torch.float64))
    fields = ("x", "y", "dx", "dy", "_one")
    J = np.zeros((n, 5, 5))
    n_split = 0
    for j, fin in enumerate(fields):
        tr = _program.Trace(torch.device("cpu"))
        seeded = dataclasses.replace(ray, **{fin: _program.FwdDual(getattr(ray, fin), 1.0)})
        val, tan = _split(_trace_model(tr.symbolize(seeded), model, FreeSpaceParaxial()), ray)
        syms = [getattr(r, f.name) for r in (val, tan) for f in dataclasses.fields(r)]
        vids = list(dict.fromkeys(v.vid for v in syms if isinstance(v, _program.Sym) and v.vid not in tr.value_slot))
        prefixes, last, prefix_comps = _program._compile_segments(tr, vids, [], (n,))
        n_split = max(n_split, len(prefixes))
        got = {}
        for comp, outs in prefix_comps:
            res = interpret(tr, comp)
            for v, t in outs.items():
                t.copy_(torch.from_numpy(res[v]))
            got.update(res)
        got.update(interpret(tr, last))
        for i, fout in enumerate(fields):
            t = getattr(tan, fout)
            J[:, i, j] = got[t.vid] if isinstance(t, _program.Sym) else t
    assert n_split >= 1
    for b in range(n):
        r = Ray(float(cols["x"][b]), float(cols["y"][b]), float(cols["dx"][b]), float(cols["dy"][b]), 0.0, 0.0)
        np.testing.assert_allclose(J[b], np.asarray(ray_matrix(model, r)), rtol=1e-12, atol=1e-15)


def test_thick_lens_matches_oracle():
    """ThickLens (reference components.py:410-451) on a scalar ray, traced for
    f32 bundles (interpreted here) and as a 5x5 matrix, against the oracle."""
    from temgymcore_amd.components import ThickLens

    model = [ThickLens(z_po=0.5, z_pi=0.6, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(64, 64))]
    omodel = [("thicklens", 0.5, 0.6, 1.0), ("detector", 1.0)]
    out = run_to_end(Ray(x=0.1, y=0.2, dx=0.3, dy=0.4, z=0.0, pathlength=0.0), model)
    ref = O.run_to_end(O.make_ray(0.1, 0.2, 0.3, 0.4), omodel)
    for k in ("x", "y", "dx", "dy", "z", "pathlength"):
        assert getattr(out, k) == ref[k], k
    x, y, dx, dy = O.plain_workload(4097, seed=11)
    ray = Ray(x=torch.from_numpy(x), y=torch.from_numpy(y), dx=torch.from_numpy(dx), dy=torch.from_numpy(dy),
              z=0.0, pathlength=0.0)
    res, _ = run_host(ray, model)
    ref = O.run_to_end(O.make_ray(x, y, dx, dy), omodel)
    for k in ("x", "y", "dx", "dy", "pathlength"):
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)
    assert res["z"] == ref["z"]
    T = ray_matrix(model, Ray(0.0, 0.0, 0.0, 0.0, 0.0, 0.0))
    np.testing.assert_allclose(T[:4, :4], [[0.6, 0, 0.7, 0], [0, 0.6, 0, 0.7], [-1, 0, 0.5, 0], [0, -1, 0, 0.5]],
                               atol=1e-15)


def test_dir_cosine_propagator_closed_form_and_device_rays():
    """FreeSpaceDirCosine on a scalar ray against the direction-cosine closed
    form (reference tests/test_rays.py:59-79, seeded distance); traced for a
    per-ray bundle it has no square-root instruction and says so."""
    import math

    from temgymcore_amd.propagator import FreeSpaceDirCosine

    d = float(np.random.default_rng(3).uniform(-10.0, 10.0))
    x0, y0, dx0, dy0, z0, pl0 = 1.0, -1.0, 2.0, 3.0, 0.5, 10.0
    out = FreeSpaceDirCosine()(Ray(x=x0, y=y0, dx=dx0, dy=dy0, z=z0, pathlength=pl0), d)
    n = math.sqrt(1.0 + dx0**2 + dy0**2)
    assert out.x == pytest.approx(x0 + (dx0 / n) / n * d, abs=1e-12)
    assert out.y == pytest.approx(y0 + (dy0 / n) / n * d, abs=1e-12)
    assert (out.dx, out.dy) == (dx0, dy0)
    assert out.z == pytest.approx(z0 + d, abs=1e-12)
    assert out.pathlength == pytest.approx(pl0 + d * n, abs=1e-12)
    # the same numbers written out: 1 + dx^2 + dy^2 = 14, so x moves by 2 d / 14
    # (the reference divides by N twice) and the path by d sqrt(14)
    assert (out.x, out.y) == (pytest.approx(1.0 + d / 7.0, abs=1e-12), pytest.approx(-1.0 + 3.0 * d / 14.0, abs=1e-12))
    assert out.pathlength == pytest.approx(10.0 + d * 14.0**0.5, abs=1e-12)
    # small slopes: the paraxial step to third order in the slope
    s = 1e-4
    par = FreeSpaceParaxial()(Ray(x=x0, y=y0, dx=s, dy=-s, z=z0, pathlength=pl0), d)
    dc = FreeSpaceDirCosine()(Ray(x=x0, y=y0, dx=s, dy=-s, z=z0, pathlength=pl0), d)
    assert abs(dc.x - par.x) <= 3 * s**3 * abs(d) and abs(dc.y - par.y) <= 3 * s**3 * abs(d)
    assert abs(dc.pathlength - par.pathlength) <= 2 * s**2 * abs(d)
    x = torch.linspace(-0.1, 0.1, 8)
    tr = _program.Trace(torch.device("cpu"))
    ray = tr.symbolize(Ray(x=x, y=x.clone(), dx=x * 0.1, dy=x * 0.2, z=0.0, pathlength=0.0))
    with pytest.raises(NotImplementedError, match="square root"):
        _trace_model(ray, [Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))], FreeSpaceDirCosine())


def test_jacobian_and_value_and_first_derivatives():
    """jacobian_and_value / calculate_derivatives (reference run.py:17-23,
    :119-147) on the README model: the README's ABCD and value."""
    from temgymcore_amd.run import calculate_derivatives, jacobian_and_value
    from temgymcore_amd.transfer import custom_jacobian_matrix

    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]
    ray = Ray(x=0.1, y=0.2, dx=0.3, dy=0.4, z=0.0, pathlength=0.0)
    jac, out = jacobian_and_value(run_to_end)(ray, model)
    readme = np.array([[.5, 0, .75, 0, 0], [0, .5, 0, .75, 0], [-1, 0, .5, 0, 0], [0, -1, 0, .5, 0], [0, 0, 0, 0, 1]])
    np.testing.assert_allclose(custom_jacobian_matrix(jac), readme, atol=1e-15)
    assert out.x == pytest.approx(0.275, abs=1e-15) and out.pathlength == pytest.approx(0.88875, abs=1e-15)
    (d1,) = calculate_derivatives(ray, model, 1)
    np.testing.assert_allclose(custom_jacobian_matrix(d1), readme, atol=1e-15)
    with pytest.raises(NotImplementedError):
        calculate_derivatives(ray, model, 2)
    with pytest.raises(NotImplementedError):
        jacobian_and_value(run_to_end, argnums=1)
