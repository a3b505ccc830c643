"""tg_stem_fit_noalloc (the descan fit with its workspace in the library's
per-GPU device array) against tg_stem_fit with a caller workspace, through
the C ABI: bitwise-equal results on two calls in a row, the second with a
smaller geometry (the array keeps the first call's contents), and a clear
error for a geometry whose workspace exceeds the array."""
import ctypes

import pytest
import torch

from temgymcore_amd import _lib, stem

pytestmark = pytest.mark.gpu


def _noalloc(model, params, targets, out, device):
    g = model.geometry()
    return _lib.load().tg_stem_fit_noalloc(ctypes.byref(g), params.data_ptr(), targets.data_ptr(), out.data_ptr(),
                                           _lib.stream_ptr(device))


def test_noalloc_matches_workspace_entry_point(device):
    lib = _lib.load()
    assert lib.tg_stem_fit_noalloc_max_bytes() == 8 << 20
    for shape, n_cam in (((64, 64), 8), ((6, 7), 3)):
        print(f"[noalloc scan={shape} n_cam={n_cam}]", flush=True)
        model, _ = stem.reference_workload(n_cam=n_cam, scan_shape=shape)
        assert lib.tg_stem_fit_workspace(ctypes.byref(model.geometry())) <= 8 << 20
        params = stem.params_tensor(model, device)
        targets = stem.stem_forward(stem.perturbed(model))["coords"]
        ref = stem.STEMFitWorkspace(model, device).run(params, targets).clone()
        out = torch.full_like(ref, float("nan"))
        _lib.check(_noalloc(model, params, targets, out, device), "tg_stem_fit_noalloc")
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (out - ref).abs().max().item()


def test_noalloc_rejects_oversized_geometry(device):
    # 128x128 scan positions x 16 camera lengths: 19 MB of workspace > 8 MiB;
    # rejected by the host check before anything is launched
    model, _ = stem.reference_workload(n_cam=16, scan_shape=(128, 128))
    assert _lib.load().tg_stem_fit_workspace(ctypes.byref(model.geometry())) > 8 << 20
    params = stem.params_tensor(model, device)
    out = torch.empty(18 + 16 + 1, dtype=torch.float64, device=device)
    dummy = torch.empty(4, dtype=torch.float32, device=device)
    st = _noalloc(model, params, dummy, out, device)
    assert st == -1
    assert "more than the" in _lib.load().tg_last_error().decode()
