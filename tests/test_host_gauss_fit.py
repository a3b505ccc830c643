"""Host-side planning of the batched Gaussian fits (no GPU needed):
validation of the instance table and the launch/workspace layout."""
import ctypes

import numpy as np
import pytest

from temgymcore_amd import _lib
from temgymcore_amd.gauss_fit import FitInstance, FitPlan, workload_params, workload_tables


def _plan(specs, n_gauss, nsigma=0.0):
    lib = _lib.load()
    arr = (FitInstance * len(specs))()
    for i, (g0, n, p0, h, w) in enumerate(specs):
        arr[i].gauss_begin, arr[i].n_gauss, arr[i].pix_begin = g0, n, p0
        arr[i].height, arr[i].width = h, w
        arr[i].pix2m[:] = [0.0, -1e-6, 0.0, 1e-6, 0.0, 0.0]
    plan = FitPlan()
    rc = lib.tg_gauss_fit_plan(arr, len(specs), n_gauss, nsigma, ctypes.byref(plan))
    return rc, plan, lib.tg_last_error().decode()


def test_plan_counts_and_layout():
    specs = [(0, 3, 0, 40, 48), (3, 0, 1920, 70, 33), (3, 600, 4230, 130, 97), (603, 2000, 16840, 64, 64)]
    rc, plan, _ = _plan(specs, 2603)
    assert rc == 0
    # slices: cdiv(n, 256) clamped to [1, 8]; tiles of 64 x 64
    nslice = [1, 1, 3, 8]
    tiles = [1 * 1, 2 * 1, 3 * 2, 1 * 1]
    assert plan.n_items == sum(a * b for a, b in zip(nslice, tiles))
    npix = [40 * 48, 70 * 33, 130 * 97, 64 * 64]
    assert plan.n_loss_blocks == sum(-(-n // 4096) for n in npix)
    assert plan.n_pix == 16840 + 4096 and plan.n_gauss == 2603 and plan.n_assigned == 2603
    assert plan.n_moment_blocks == -(-2603 // 4)
    assert plan.cut == 80.0
    offs = [plan.off_inst, plan.off_items, plan.off_loss_blocks, plan.off_order, plan.off_gauss_inst, plan.off_coef,
            plan.off_unit, plan.off_partial, plan.off_weight, plan.off_loss_part, plan.off_moments]
    assert offs == sorted(offs) and all(o % 256 == 0 for o in offs)
    assert plan.off_partial - plan.off_unit >= 136 * 2603 and plan.off_unit - plan.off_coef >= 136 * 2603
    assert plan.off_weight - plan.off_partial >= 8 * sum(s * n for s, n in zip(nslice, npix))
    assert plan.off_loss_part - plan.off_weight >= 8 * (plan.n_pix + 16)  # tail padding of the weights
    assert plan.workspace_bytes >= plan.off_moments + 96 * 2603
    assert plan.table_bytes == plan.off_coef


def test_plan_cutoff():
    rc, plan, _ = _plan([(0, 1, 0, 8, 8)], 1, nsigma=4.0)
    assert rc == 0 and plan.cut == 16.0


@pytest.mark.parametrize("specs, n_gauss, msg", [
    ([(0, 3, 0, 8, 8), (2, 3, 64, 8, 8)], 5, "share Gaussian rows"),
    ([(0, 3, 0, 8, 8), (3, 3, 63, 8, 8)], 6, "share pixels"),
    ([(0, 4, 0, 8, 8)], 3, "outside"),
    ([(0, 1, 0, 0, 8)], 1, "empty image"),
    ([(0, 1, -1, 8, 8)], 1, "negative pixel offset"),
])
def test_plan_rejects_bad_tables(specs, n_gauss, msg):
    rc, _, err = _plan(specs, n_gauss)
    assert rc == -1 and msg in err


def test_plan_allows_unassigned_rows_and_empty_instances():
    rc, plan, _ = _plan([(2, 0, 0, 8, 8), (5, 2, 64, 8, 8)], 10)
    assert rc == 0 and plan.n_assigned == 2 and plan.n_gauss == 10


def test_workload_tables_are_seeded_and_in_range():
    counts, sides = workload_tables()
    assert len(counts) == 256 and counts.min() >= 3 and counts.max() <= 2000
    assert sides.min() >= 128 and sides.max() <= 512
    c2, s2 = workload_tables()
    assert np.array_equal(counts, c2) and np.array_equal(sides, s2)
    p = workload_params(counts[:4], sides[:4])
    assert p.shape == (counts[:4].sum(), 16) and np.all(np.isinf(p[:, 9:11]))
    half = 0.5 * sides[0] * 1e-6
    assert np.abs(p[:counts[0], 0:2]).max() <= 0.8 * half
