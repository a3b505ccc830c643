"""The plan cache behind the drop-in calls (run_to_end, run_to_image,
Grid.into_image, component and propagator calls on tensor rays), on the host:
what its key distinguishes, and that a template bound to a new call's arrays
equals a fresh compile of that call byte for byte, except for the output
buffers (fresh allocations in both)."""
import numpy as np
import pytest
import torch

from oracle import ray_oracle as O
from temgymcore_amd import _lib, _program
from temgymcore_amd.components import Deflector, Detector, Lens
from temgymcore_amd.grid import make_bin_spec
from temgymcore_amd.propagator import FreeSpaceParaxial
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import _trace_model


def _ray(n=1000, seed=0, dtype=np.float32):
    x, y, dx, dy = O.plain_workload(n, seed=seed)
    t = [torch.from_numpy(a.astype(dtype)) for a in (x, y, dx, dy)]
    return Ray(x=t[0], y=t[1], dx=t[2], dy=t[3], z=0.0, pathlength=0.0)


def _model(f=1.0):
    return [Lens(z=0.5, focal_length=f), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]


def _key(*objs):
    leaves = []
    return _program._structure_key(objs, leaves, {}), leaves


def test_key_ignores_array_identity_and_values():
    k1, l1 = _key("run_to_end", tuple(_model()), FreeSpaceParaxial(), _ray(seed=0))
    k2, l2 = _key("run_to_end", tuple(_model()), FreeSpaceParaxial(), _ray(seed=1))
    assert k1 == k2 and hash(k1) == hash(k2)
    assert len(l1) == len(l2) == 4 and all(a is not b for a, b in zip(l1, l2))


@pytest.mark.parametrize("change", ["focal_length", "dtype", "shape", "stride", "scalar_type", "scalar_value",
                                    "alias", "propagator_state", "element"])
def test_key_distinguishes(change):
    base = ("run_to_end", tuple(_model()), FreeSpaceParaxial(), _ray())
    r = _ray()
    model = tuple(_model())
    prop = FreeSpaceParaxial()
    if change == "focal_length":
        model = tuple(_model(f=1.0 + 2 ** -40))
    elif change == "dtype":
        r = _ray(dtype=np.float64)
    elif change == "shape":
        r = _ray(n=1001)
    elif change == "stride":
        big = _ray(n=2000)
        r = Ray(x=big.x[::2], y=big.y[::2], dx=big.dx[::2], dy=big.dy[::2], z=0.0, pathlength=0.0)
    elif change == "scalar_type":
        r = Ray(x=r.x, y=r.y, dx=r.dx, dy=r.dy, z=np.float32(0.0), pathlength=0.0)
    elif change == "scalar_value":
        r = Ray(x=r.x, y=r.y, dx=r.dx, dy=r.dy, z=-0.0, pathlength=0.0)
    elif change == "alias":
        r = Ray(x=r.x, y=r.x, dx=r.dx, dy=r.dy, z=0.0, pathlength=0.0)
    elif change == "propagator_state":
        prop = FreeSpaceParaxial()
        prop.extra = 1.0
    elif change == "element":
        model = (Deflector(z=0.5, def_x=0.0, def_y=0.0), model[1])
    assert _key(*base)[0] != _key("run_to_end", model, prop, r)[0]


def test_key_one_element_arrays_by_value_and_uncacheable_objects():
    a = _key(Lens(z=torch.tensor([0.5]), focal_length=1.0))[0]
    b = _key(Lens(z=torch.tensor([0.25]), focal_length=1.0))[0]
    assert a != b
    with pytest.raises(_program.Uncacheable):
        _key(Lens(z=0.5, focal_length=object()))
    state = 1.0
    with pytest.raises(_program.Uncacheable):
        _key(lambda r: r.x * state)  # a closure may change behind the key


def _cpu_plan(ray, model, acc=None):
    tr = _program.Trace(torch.device("cpu"))
    out = _trace_model(tr.symbolize(ray), model, FreeSpaceParaxial())
    req = None
    if acc is not None:
        grid = model[-1]
        req = (make_bin_spec(grid.metres_to_pixels_mat, grid.shape, acc), out.x, out.y)
    return _program.build_plan(tr, out, req)


def _masked(prog):
    p = _lib.Program.from_buffer_copy(prog)
    for k in range(p.n_slot):
        p.slot[k].ptr = 0
    for k in range(p.n_out):
        p.out[k] = 0
    return bytes(p)


def _long_model():
    model = []
    for i in range(14):
        z = 0.05 * (i + 1)
        model.append(Lens(z=z, focal_length=0.7 + 0.01 * i) if i % 2 else Deflector(z=z, def_x=1e-3 * i, def_y=-2e-4))
    return model + [Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(8, 8))]


@pytest.mark.parametrize("which", ["plain", "binning", "split"])
def test_bound_template_equals_fresh_compile(which):
    model = _long_model() if which == "split" else _model()
    acc1 = acc2 = None
    if which == "binning":
        acc1, acc2 = torch.zeros((128, 128), dtype=torch.int32), torch.zeros((128, 128), dtype=torch.int32)
    r1, r2 = _ray(seed=0), _ray(seed=3)
    k1, leaves1 = _key("t", tuple(model), r1, acc1)
    k2, leaves2 = _key("t", tuple(model), r2, acc2)
    assert k1 == k2
    plan1 = _cpu_plan(r1, model, acc1)
    tpl = _program._PlanTemplate(plan1, leaves1)
    bound, out_ray, (inputs, outs) = tpl.bind(leaves2, acc2)
    fresh = _cpu_plan(r2, model, acc2)
    assert len(bound) == len(fresh.programs) == len(plan1.programs)
    if which == "split":
        assert len(bound) >= 2
    out_ptr = {}
    for (prog, spec), (fprog, fspec, _), (_, _, binds, groups) in zip(bound, fresh.programs, tpl.programs):
        assert _masked(prog) == _masked(fprog)
        for k, (kind, x) in enumerate(binds):
            if kind == "leaf":
                # the same input tensors: the same addresses as the fresh compile
                assert prog.slot[k].ptr == fprog.slot[k].ptr == inputs[x].data_ptr()
                assert inputs[x] is leaves2[x]
            else:
                # an intermediate buffer written by an earlier program of this call
                assert prog.slot[k].ptr == out_ptr[x]
        n_out = 0
        for (shape, dt), members in groups:
            for k, vid in members:
                assert prog.out[k] == outs[vid].data_ptr() and outs[vid].shape == shape and outs[vid].dtype == dt
                out_ptr[vid] = prog.out[k]
                n_out += 1
        assert n_out == prog.n_out == fprog.n_out
        assert len(set(out_ptr.values())) == len(out_ptr)  # distinct buffers
        if spec is None:
            assert fspec is None
        else:
            assert spec.acc == acc2.data_ptr() != acc1.data_ptr()
            a, b = _lib.BinSpec.from_buffer_copy(spec), _lib.BinSpec.from_buffer_copy(fspec)
            assert bytes(a) == bytes(b)
    for name in ("x", "y", "dx", "dy", "pathlength"):
        v = getattr(out_ray, name)
        assert isinstance(v, torch.Tensor) and v.shape == r2.x.shape
        assert all(v is not t for t in leaves1)
    assert out_ray.z == fresh.out_ray.z


def test_passthrough_input_field_is_the_new_call_array():
    # no element rewrites _one (propagators and components only read it):
    # the output ray returns the caller's own _one array
    model = [Deflector(z=0.0, def_x=1e-3, def_y=0.0)]

    def ray(seed):
        r = _ray(seed=seed)
        return Ray(x=r.x, y=r.y, dx=r.dx, dy=r.dy, z=0.0, pathlength=0.0, _one=torch.ones(1000))

    r1, r2 = ray(0), ray(1)
    k1, leaves1 = _key("t", tuple(model), r1)
    k2, leaves2 = _key("t", tuple(model), r2)
    assert k1 == k2
    tpl = _program._PlanTemplate(_cpu_plan(r1, model), leaves1)
    _, out_ray, _ = tpl.bind(leaves2, None)
    assert out_ray._one is r2._one


def test_cache_lru_and_uncacheable_paths(monkeypatch):
    cache = _program.PlanCache(max_entries=2)
    launched = []

    class FakePlan:
        def __init__(self, tag):
            self.tag = tag

        def launch(self, stream=None):
            launched.append(self.tag)
            return self.tag

    # templates need real plans; here only the bookkeeping is checked
    monkeypatch.setattr(_program, "_PlanTemplate", lambda plan, leaves: (_ for _ in ()).throw(_program.Uncacheable()))
    assert cache.launch(("a", 1.0), lambda: FakePlan("a")) == "a"
    assert cache.misses == 1 and len(cache) == 1
    # a key whose plan could not be templated is remembered as uncacheable
    assert cache.launch(("a", 1.0), lambda: FakePlan("a2")) == "a2"
    assert cache.uncached == 1 and cache.hits == 0
    cache.launch(("b", 1.0), lambda: FakePlan("b"))
    cache.launch(("c", 1.0), lambda: FakePlan("c"))
    assert len(cache) == 2
    # an unkeyable call always builds
    cache.launch(("d", object()), lambda: FakePlan("d"))
    assert cache.uncached == 2 and launched[-1] == "d"
