"""Argument validation of C-ABI entry points, exercised without a GPU: each
bad argument is rejected with TG_EINVAL (-1) and a message naming it before
any HIP call is made (the pointers passed are never dereferenced)."""
import ctypes

import pytest

# each module declares its entry points' ctypes signatures when imported
from temgymcore_amd import _lib, gauss_fit, gaussian, source, stem  # noqa: F401

BAD = ctypes.c_void_p(16)  # non-NULL, never dereferenced on these paths


def _err(lib):
    return lib.tg_last_error().decode()


@pytest.mark.parametrize("kw, msg", [
    (dict(n=-1), "n < 0"),
    (dict(kind=2), "kind"),
    (dict(dtype=7), "dtype"),
    (dict(n_rings=0), "empty ring"),
    (dict(x=None), "NULL"),
])
def test_source_rings_rejects(kw, msg):
    lib = _lib.load()
    a = dict(radius=BAD, step=BAD, ring_first=BAD, n_rings=1, seg_first=BAD, n_seg=1, n=10, kind=0, off_x=0.0,
             off_y=0.0, dtype=_lib.TG_F64, x=BAD, y=BAD, dx=BAD, dy=BAD, stream=None)
    a.update(kw)
    rc = lib.tg_source_rings(*a.values())
    assert rc == -1 and msg in _err(lib)


def test_source_rings_empty_is_a_no_op():
    lib = _lib.load()
    rc = lib.tg_source_rings(None, None, None, 0, None, 0, 0, 0, 0.0, 0.0, _lib.TG_F64, None, None, None, None, None)
    assert rc == 0


@pytest.mark.parametrize("args, msg", [
    ((-1, 1.0, 0.0, 1.0, 1.0, 1, BAD, BAD, None), "n < 0"),
    ((5, 1.0, 0.0, 1.0, 1.0, 9, BAD, BAD, None), "dtype"),
    ((5, 1.0, 0.0, 1.0, 1.0, 1, None, BAD, None), "NULL"),
])
def test_source_fibonacci_rejects(args, msg):
    lib = _lib.load()
    assert lib.tg_source_fibonacci(*args) == -1 and msg in _err(lib)


@pytest.mark.parametrize("args, msg", [
    ((-2, 1, 1.0, 0, 0.0, 0.0, 1, BAD, BAD, BAD, BAD, None), "n < 0"),
    ((4, 1, 1.0, 3, 0.0, 0.0, 1, BAD, BAD, BAD, BAD, None), "kind"),
    ((4, 1, 1.0, 0, 0.0, 0.0, 1, BAD, None, BAD, BAD, None), "NULL"),
])
def test_source_disc_rejects(args, msg):
    lib = _lib.load()
    assert lib.tg_source_disc(*args) == -1 and msg in _err(lib)


def _geometry(**kw):
    model, _ = stem.reference_workload(n_cam=2, scan_shape=(4, 4), n_side=8)
    g = model.geometry()
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def test_backproject_rejects():
    lib = _lib.load()
    g = _geometry()
    assert lib.tg_stem_backproject(ctypes.byref(g), BAD, BAD, 2, BAD, None, None) == -1
    assert "camera 2 out of range" in _err(lib)
    assert lib.tg_stem_backproject(ctypes.byref(g), BAD, None, 0, BAD, None, None) == -1
    assert "NULL" in _err(lib)
    g = _geometry(n_side=7)
    assert lib.tg_stem_backproject(ctypes.byref(g), BAD, BAD, 0, BAD, None, None) == -1
    assert "n_side" in _err(lib)
    # one scan position's detector image is addressed by 32-bit byte offsets
    g = _geometry(det_h=1 << 15, det_w=(1 << 15) + 1)
    assert lib.tg_stem_backproject(ctypes.byref(g), BAD, BAD, 0, BAD, None, None) == -1
    assert "exceeds 2^30" in _err(lib)


def test_fit_terms_and_fit_reject():
    lib = _lib.load()
    g = _geometry()
    assert lib.tg_stem_fit_terms(ctypes.byref(g), BAD, None, BAD, None) == -1 and "NULL" in _err(lib)
    misaligned = ctypes.c_void_p(24)
    assert lib.tg_stem_fit_terms(ctypes.byref(g), BAD, misaligned, BAD, None) == -1 and "aligned" in _err(lib)
    assert lib.tg_stem_fit(ctypes.byref(g), BAD, BAD, BAD, None, None) == -1 and "NULL" in _err(lib)
    g = _geometry(n_cam=0)
    assert lib.tg_stem_fit(ctypes.byref(g), BAD, BAD, BAD, BAD, None) == -1 and "n_cam" in _err(lib)


def test_fit_workspace_size():
    lib = _lib.load()
    g = _geometry()
    S, np_, C = 16, 18 + 2 + 1, 2
    assert lib.tg_stem_fit_workspace(ctypes.byref(g)) == 8 * S * (np_ + C * 7)
    assert lib.tg_stem_fit_workspace(None) == 0


def test_stem_geometry_cache_copies_and_tracks_fields():
    """model.geometry() is cached on the model: every call returns a fresh
    copy equal to a rebuild, and a change of any field it depends on is seen."""
    model, _ = stem.reference_workload(n_cam=2, scan_shape=(4, 4), n_side=8)
    g = model.geometry()
    assert bytes(g) == bytes(model._build_geometry())
    g.n_side = 7  # mutating a returned struct leaves the cache intact
    assert model.geometry().n_side == 8
    for field, value in [("det_pixel", 3e-6), ("camera_lengths", (0.2, 0.5, 0.9)), ("defocus", 150e-9),
                         ("sample_pixel", 3e-11), ("descan_offsets", (1e-9, 0.0, 0.0, 2e-6))]:
        before = bytes(model.geometry())
        setattr(model, field, value)
        after = model.geometry()
        assert bytes(after) == bytes(model._build_geometry()) and bytes(after) != before, field
