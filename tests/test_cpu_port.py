"""The OpenMP C port timed as the CPU baseline agrees with the oracle."""
import numpy as np
import pytest

from oracle import c as cport
from oracle import ray_oracle as O
from oracle import stem_oracle as SO


@pytest.fixture(scope="module")
def lib():
    try:
        return cport.load()
    except Exception as e:  # no C compiler
        pytest.skip(str(e))


def test_plain_port_bit_exact(lib):
    n = 100_003
    x, y, dx, dy = O.plain_workload(n, seed=1)
    outs = [np.empty(n, np.float32) for _ in range(5)]
    hist = np.zeros((128, 128), np.int64)
    M = O.metres_to_pixels_matrix(**O.PLAIN_DET)
    m = np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 0], M[1, 1], M[1, 2]])
    lib.plain_run(n, x.ctypes.data, y.ctypes.data, dx.ctypes.data, dy.ctypes.data, 1.0, m.ctypes.data, 128, 128,
                  *[o.ctypes.data for o in outs], hist.ctypes.data)
    ref, h = O.plain_run(x, y, dx, dy)
    for o, k in zip(outs, ("x", "y", "dx", "dy", "pathlength")):
        np.testing.assert_array_equal(o, ref[k])
    np.testing.assert_array_equal(hist, h)


def test_stem_fit_port_vs_fd(lib):
    geom = dict(scan_shape=(4, 5), n_side=8, semi_conv=0.5e-3, defocus=2e-7, pivot=1e4, offsets=(0.0,) * 4)
    rng = np.random.default_rng(4)
    p = np.concatenate([rng.normal(0, 1e-3, 16), [1e-10, 12.0], [0.2, 0.9]])
    q = p * (1 + 1e-3 * rng.normal(size=p.shape))
    tx, ty = SO.forward(geom, q)
    targets = np.ascontiguousarray(np.stack([tx, ty], axis=-1).astype(np.float32))
    out = np.zeros(21)
    off = np.zeros(4)
    lib.stem_fit(4, 5, 8, 2, 0.5e-3, 2e-7, 1e4, off.ctypes.data, p.ctypes.data, targets.ctypes.data, 0, 20,
                 out.ctypes.data)
    g = SO.fd_grad(geom, p, targets)
    assert out[-1] == pytest.approx(SO.loss(geom, p, targets), rel=1e-9)
    for blk in (slice(0, 16), slice(16, 17), slice(17, 18), slice(18, 20)):
        np.testing.assert_allclose(out[:-1][blk], g[blk], rtol=0, atol=1e-5 * np.abs(g[blk]).max())


def test_stem_fit_port_gradient_at_zero_scan_step(lib):
    # d(scan position)/d(step) is the rotated integer pixel offset, finite at
    # step 0 (a quotient sp / step would be 0/0 there)
    geom = dict(scan_shape=(4, 5), n_side=8, semi_conv=0.5e-3, defocus=2e-7, pivot=1e4, offsets=(0.0,) * 4)
    rng = np.random.default_rng(6)
    p = np.concatenate([rng.normal(0, 1e-3, 16), [0.0, 12.0], [0.2, 0.9]])
    q = p.copy()
    q[16] = 1e-10
    tx, ty = SO.forward(geom, q)
    targets = np.ascontiguousarray(np.stack([tx, ty], axis=-1).astype(np.float32))
    out = np.zeros(21)
    off = np.zeros(4)
    lib.stem_fit(4, 5, 8, 2, 0.5e-3, 2e-7, 1e4, off.ctypes.data, p.ctypes.data, targets.ctypes.data, 0, 20,
                 out.ctypes.data)
    steps = SO.fd_steps(p)
    steps[16] = 1e-10
    g = SO.fd_grad(geom, p, targets, steps=steps)
    assert np.all(np.isfinite(out))
    assert abs(g[16]) > 0
    for blk in (slice(0, 16), slice(16, 17), slice(17, 18), slice(18, 20)):
        np.testing.assert_allclose(out[:-1][blk], g[blk], rtol=0, atol=1e-5 * np.abs(g[blk]).max())
