"""The plan cache on the GPU: repeated drop-in calls (run_to_end,
run_to_image, Grid.into_image, component calls) re-launch cached programs on
new arrays, bit-exact against the oracle; a changed parameter, shape or dtype
misses the cache; and a cached call costs little host time."""
import time

import numpy as np
import pytest
import torch

from oracle import ray_oracle as O
from temgymcore_amd import _program
from temgymcore_amd.components import Deflector, Detector, Lens
from temgymcore_amd.ray import Ray
from temgymcore_amd.run import run_to_end, run_to_image

pytestmark = pytest.mark.gpu

DET = dict(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))


def _ray(device, n=1 << 20, seed=0, dtype=np.float32):
    x, y, dx, dy = O.plain_workload(n, seed=seed, dtype=dtype)
    t = [torch.from_numpy(a).to(device) for a in (x, y, dx, dy)]
    return Ray(x=t[0], y=t[1], dx=t[2], dy=t[3], z=0.0, pathlength=0.0), (x, y, dx, dy)


def _model(f=1.0):
    return [Lens(z=0.5, focal_length=f), Detector(**DET)]


def _check(out, host, f=1.0):
    ref = O.run_to_end(O.make_ray(*host), [("lens", 0.5, f), ("detector", 1.0)])
    for k in ("x", "y", "dx", "dy", "pathlength"):
        np.testing.assert_array_equal(getattr(out, k).cpu().numpy(), ref[k], err_msg=k)
    return ref


def test_cached_run_to_end_hits_and_misses(device):
    cache = _program.plan_cache
    cache.clear()
    r0, h0 = _ray(device, seed=0)
    _check(run_to_end(r0, _model()), h0)
    assert (cache.misses, cache.hits) == (1, 0)
    # new arrays, same structure: a hit, on the new data
    for seed in (1, 2):
        r, h = _ray(device, seed=seed)
        _check(run_to_end(r, _model()), h)
    assert (cache.misses, cache.hits) == (1, 2)
    # a changed parameter value, ray count or dtype misses
    r, h = _ray(device, seed=3)
    _check(run_to_end(r, _model(f=0.75)), h, f=0.75)
    r, h = _ray(device, n=1000, seed=4)
    _check(run_to_end(r, _model()), h)
    r, h = _ray(device, seed=5, dtype=np.float64)
    _check(run_to_end(r, _model()), h)
    assert (cache.misses, cache.hits) == (4, 2)
    # outputs are fresh buffers every call
    a = run_to_end(r0, _model())
    b = run_to_end(r0, _model())
    assert a.x.data_ptr() != b.x.data_ptr()
    torch.testing.assert_close(a.x, b.x, rtol=0, atol=0)


def test_cached_run_to_image_and_into_image(device):
    cache = _program.plan_cache
    cache.clear()
    for seed in (0, 1, 2):
        r, h = _ray(device, seed=seed)
        out, img = run_to_image(r, _model())
        _, hist = O.plain_run(*h)
        _check(out, h)
        np.testing.assert_array_equal(img.cpu().numpy().astype(np.int64), hist)
        # the same positions binned by Grid.into_image (cached separately)
        acc = torch.zeros((128, 128), dtype=torch.int64, device=device)
        _model()[-1].into_image(out, acc)
        np.testing.assert_array_equal(acc.cpu().numpy(), hist)
    assert cache.hits == 4 and cache.misses == 2


def test_cached_component_calls_in_run_iter_style(device):
    cache = _program.plan_cache
    cache.clear()
    defl = Deflector(z=0.0, def_x=1e-3, def_y=-2e-3)
    for seed in (0, 1):
        r, (x, y, dx, dy) = _ray(device, n=4096, seed=seed)
        out = defl(r)
        ref = O.apply(O.make_ray(x, y, dx, dy), ("deflector", 0.0, 1e-3, -2e-3))
        for k in ("x", "y", "dx", "dy", "pathlength"):
            np.testing.assert_array_equal(getattr(out, k).cpu().numpy(), np.asarray(ref[k], np.float32), err_msg=k)
    assert cache.hits == 1


def test_cached_call_host_time(device):
    """A repeated run_to_end on 1M rays (Lens + Detector) costs tens of
    microseconds of host time; tracing and compiling cost ~1 ms."""
    r, _ = _ray(device)
    model = _model()
    run_to_end(r, model)
    torch.cuda.synchronize()
    per_call = []
    for _ in range(5):
        t0 = time.perf_counter()
        for _ in range(100):
            run_to_end(r, model)
        per_call.append((time.perf_counter() - t0) / 100)
        torch.cuda.synchronize()
    _program.plan_cache.clear()
    t0 = time.perf_counter()
    for _ in range(20):
        _program.plan_fn(lambda q: q, r)  # trace + compile of a trivial model, for scale
    uncached = (time.perf_counter() - t0) / 20
    host_us = 1e6 * min(per_call)
    print(f"cached run_to_end host time {host_us:.1f} us per call (trivial uncached trace {1e6 * uncached:.0f} us)",
          flush=True)
    assert host_us < 80.0
