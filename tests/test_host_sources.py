"""Ray sources on the host (no GPU): the oracle's concentric rings pinned by a
hand-worked partitioned sum and their geometry, and the ring table that the
device generator walks (replayed here point by point in numpy) against the
oracle's sequential loop, bit for bit."""
import numpy as np
import pytest

from oracle import ray_oracle as O
from temgymcore_amd import source as S


def test_multi_cumsum_known_answer():
    """Partitions [2, 3] of [a, a, b, b, b]: the counter restarts one point
    late (reference utils.py:46-80), giving [0, a, a + b, 0, b]."""
    a, b = 0.25, 0.5
    v = np.array([a, a, b, b, b])
    O.multi_cumsum_inplace(v, np.array([2, 3]), 0.0)
    np.testing.assert_array_equal(v, [0.0, a, a + b, 0.0, b])
    # partitions [3, 4, 2] of [a, a, a, b, b, b, b, c, c]: restarts at 4 and 9
    c = 1.0
    w = np.array([a, a, a, b, b, b, b, c, c])
    O.multi_cumsum_inplace(w, np.array([3, 4, 2]), 0.0)
    np.testing.assert_array_equal(w, [0.0, a, 2 * a, 2 * a + b, 0.0, b, 2 * b, 2 * b + c, 2 * b + 2 * c])


@pytest.mark.parametrize("num", [1, 7, 50, 4096, 20000])
def test_oracle_rings_geometry(num):
    """Points lie on radii linspace(0, R, n + 1)[1:], ring k holding about
    2 pi k num / sum points; the first point of the first ring is at angle 0."""
    R = 2.5e-3
    r, a = O.ring_params(num, R)
    rings = np.unique(r)
    n_rings = max(1, int(np.floor((-1 + np.sqrt(1 + 4 * num / np.pi)) / 2)))
    np.testing.assert_array_equal(rings, np.linspace(0, R, n_rings + 1)[1:][: rings.size])
    assert abs(r.size - num) <= max(3, 0.02 * num)
    assert a[0] == 0.0
    yx = O.concentric_rings(num, R)
    np.testing.assert_allclose(np.hypot(yx[:, 0], yx[:, 1]), r, rtol=1e-15)


@pytest.mark.parametrize("num", [1, 3, 7, 50, 333, 4096, 20000])
def test_ring_table_replay_matches_sequential_angles(num):
    """Replaying the ring table the way the device kernel does (each point
    sums, in order, the steps of the points before it in its restart run)
    gives the oracle's angles and radii bit for bit."""
    R = 0.5e-3
    r_ref, a_ref = O.ring_params(num, R)
    radii, steps, ring_first, seg_first, n = S.ring_table(num, R)
    assert n == r_ref.size
    a = np.empty(n)
    r = np.empty(n)
    for v in range(n):
        s = seg_first[np.searchsorted(seg_first, v, side="right") - 1]
        acc = 0.0
        for u in range(s + 1, v + 1):
            acc = acc + steps[np.searchsorted(ring_first, u, side="right") - 1]
        a[v] = acc
        r[v] = radii[np.searchsorted(ring_first, v, side="right") - 1]
    np.testing.assert_array_equal(a, a_ref)
    np.testing.assert_array_equal(r, r_ref)


def test_host_generators_match_oracle():
    for num in (50, 4096):
        np.testing.assert_array_equal(S.concentric_rings(num, 1e-3), O.concentric_rings(num, 1e-3))
        np.testing.assert_array_equal(S.PointSource(z=0.0, semi_conv=1e-3, offset_xy=(1e-6, -2e-6)).generate_array(num),
                                      O.point_source_array(num, 1e-3, (1e-6, -2e-6)))
        np.testing.assert_array_equal(S.ParallelBeam(z=0.0, radius=1e-3, offset_xy=(1e-6, -2e-6)).generate_array(num),
                                      O.parallel_beam_array(num, 1e-3, (1e-6, -2e-6)))


def test_cpu_make_rays_stays_on_host():
    import torch

    rays = S.PointSource(z=0.1, semi_conv=1e-3).make_rays(100, device="cpu")
    assert rays.x.device.type == "cpu" and rays.dx.numel() == O.ring_params(100, 1e-3)[0].size
    one = S.PointSource(z=0.1, semi_conv=1e-3).make_rays(1, device="cpu")
    assert isinstance(one.dx, float)
    del torch
