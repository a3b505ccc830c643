"""The library's tuning table and its consumers, on the host (no GPU)."""
import ctypes

import pytest

from temgymcore_amd import _lib, autotune
from temgymcore_amd.gauss_fit import FitInstance, FitPlan


@pytest.fixture(autouse=True)
def _clean():
    autotune.clear()
    yield
    autotune.clear()


def test_set_get_clear():
    assert autotune.get_choice(autotune.BEAMLET_SPLITS, (10, 20, 30)) is None
    autotune.set_choice(autotune.BEAMLET_SPLITS, (10, 20, 30), 7)
    assert autotune.get_choice(autotune.BEAMLET_SPLITS, (10, 20, 30)) == 7
    assert autotune.get_choice(autotune.BEAMLET_SPLITS, (10, 20, 31)) is None
    assert autotune.get_choice(autotune.RAY_GRID, (10, 20, 30)) is None
    autotune.clear()
    assert autotune.get_choice(autotune.BEAMLET_SPLITS, (10, 20, 30)) is None


def test_rejects_bad_entries():
    with pytest.raises(ValueError):
        autotune.set_choice(99, (1, 2, 3), 4)
    with pytest.raises(ValueError):
        autotune.set_choice(autotune.RAY_GRID, (1, 2, 3), 0)


def test_beamlet_workspace_follows_tuned_splits():
    lib = _lib.load()
    n, H, W = 100000, 1024, 1024
    # after the slices: the per-launch tile-sum path flags, then 96 bytes of
    # separable-factor parameters per beamlet
    extra = 256 * 4 + 96 * n
    autotune.set_choice(autotune.BEAMLET_SPLITS, (n, H, W), 16)
    assert lib.tg_beamlet_sum_workspace(n, H, W) == 16 * H * W * 8 + extra
    autotune.set_choice(autotune.BEAMLET_SPLITS, (n, H, W), 1000)  # clamped to 128
    assert lib.tg_beamlet_sum_workspace(n, H, W) == 128 * H * W * 8 + extra
    # slices are padded to 256 bytes before the flags
    assert lib.tg_beamlet_sum_workspace(3, 3, 5) == 256 + 256 * 4 + 96 * 3


def test_beamlet_path_knob_is_accepted():
    key = (1000, 64, 64)
    autotune.set_choice(autotune.BEAMLET_PATH, key, autotune.BEAMLET_PATH_VALU)
    assert autotune.get_choice(autotune.BEAMLET_PATH, key) == autotune.BEAMLET_PATH_VALU
    autotune.clear()
    assert autotune.get_choice(autotune.BEAMLET_PATH, key) is None


def test_tile_shape_knob_checks_values():
    key = (1000, 64, 64)
    for shape in autotune.TILE_SHAPES:
        autotune.set_choice(autotune.TILE_SHAPE, key, shape)
        assert autotune.get_choice(autotune.TILE_SHAPE, key) == shape
    with pytest.raises(ValueError):
        autotune.set_choice(autotune.TILE_SHAPE, key, len(autotune.TILE_SHAPES) + 1)
    with pytest.raises(ValueError):
        autotune.set_choice(autotune.BEAMLET_PATH, key, 3)


def test_fit_plan_records_tuned_tile_shape():
    """The plan's forward items tile each instance with the tuned shape."""
    lib = _lib.load()
    arr = (FitInstance * 2)()
    dims = [(150, 130), (40, 300)]  # (height, width)
    for i, (h, w) in enumerate(dims):
        arr[i].gauss_begin, arr[i].n_gauss, arr[i].pix_begin = 10 * i, 10, sum(a * b for a, b in dims[:i])
        arr[i].height, arr[i].width = h, w
        arr[i].pix2m[:] = [0.0, -1e-6, 0.0, 1e-6, 0.0, 0.0]
    npix = sum(a * b for a, b in dims)

    def plan():
        p = FitPlan()
        assert lib.tg_gauss_fit_plan(arr, 2, 20, 0.0, ctypes.byref(p)) == 0
        return p

    p = plan()
    assert p.tile_shape == autotune.TILE_64x64_RUN16
    for shape, (tw, th, _) in autotune.TILE_SHAPES.items():
        autotune.set_choice(autotune.TILE_SHAPE, (2, 20, npix, 1), shape)
        p = plan()
        assert p.tile_shape == shape
        tiles = sum(-(-h // th) * -(-w // tw) for h, w in dims)
        assert p.n_items == tiles  # one slice per instance (10 Gaussians)


def test_fit_plan_follows_tuned_slice():
    lib = _lib.load()
    arr = (FitInstance * 1)()
    arr[0].gauss_begin, arr[0].n_gauss, arr[0].pix_begin, arr[0].height, arr[0].width = 0, 2000, 0, 64, 64
    arr[0].pix2m[:] = [0.0, -1e-6, 0.0, 1e-6, 0.0, 0.0]

    def items():
        plan = FitPlan()
        assert lib.tg_gauss_fit_plan(arr, 1, 2000, 0.0, ctypes.byref(plan)) == 0
        return plan.n_items

    assert items() == 8  # default 256 Gaussians per slice, at most 8 slices
    autotune.set_choice(autotune.FIT_SLICE, (1, 2000, 4096), 1000)
    assert items() == 2
    autotune.set_choice(autotune.FIT_SLICE, (1, 2000, 4096), 4096)
    assert items() == 1


def test_save_and_load(tmp_path):
    autotune.set_choice(autotune.RAY_GRID, (5, 6, 0), 512)
    path = tmp_path / "tune.json"
    autotune.save(str(path), [(autotune.RAY_GRID, (5, 6, 0), 512)])
    autotune.clear()
    assert autotune.load(str(path)) == 1
    assert autotune.get_choice(autotune.RAY_GRID, (5, 6, 0)) == 512
