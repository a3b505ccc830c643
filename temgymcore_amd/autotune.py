"""Autotuner of launch shapes.

Each tuner times a few candidate launch choices of one kernel on the real
inputs (HIP events on torch's current stream, median of `reps` runs after a
warm-up) and records the fastest in the library's tuning table
(``tg_tune_set``), which the entry points consult on every later call with
the same problem sizes.  Tuning is explicit: nothing here runs unless called,
and ``clear()`` restores the built-in heuristics.

Knobs (include/temgym_hip.h, ``tg_tune_knob``):
  BEAMLET_SPLITS  beamlet slices per detector tile of ``tg_beamlet_sum``
  RAY_GRID        workgroups of a binning ``tg_ray_program_run``
  FIT_SLICE       Gaussians per forward slice of the batched fits
  STEM_FIT        launch variant of the descan fit (per scan position, or
                  per scan position and camera length)
  BEAMLET_PATH    tile sum of ``tg_beamlet_sum``: the matrix-core product for
                  separable beamlets (when all are), or the VALU recurrence
  TILE_SHAPE      tile shape and run length of the VALU recurrence tile sum
                  (``tg_beamlet_sum`` and the batched fits' forward pass)
"""
from __future__ import annotations

import ctypes
import json

import torch

from . import _lib

BEAMLET_SPLITS, RAY_GRID, FIT_SLICE, STEM_FIT, BEAMLET_PATH, TILE_SHAPE = 1, 2, 3, 4, 5, 6
_NAMES = {BEAMLET_SPLITS: "beamlet_splits", RAY_GRID: "ray_grid", FIT_SLICE: "fit_slice", STEM_FIT: "stem_fit",
          BEAMLET_PATH: "beamlet_path", TILE_SHAPE: "tile_shape"}
STEM_FIT_SCAN, STEM_FIT_CAMERA = 1, 2
BEAMLET_PATH_AUTO, BEAMLET_PATH_VALU = 1, 2
# tg_tile_shape: (width, height, run length) of each tile shape
TILE_64x64_RUN16, TILE_128x32_RUN16, TILE_32x128_RUN16, TILE_64x32_RUN8 = 1, 2, 3, 4
TILE_SHAPES = {TILE_64x64_RUN16: (64, 64, 16), TILE_128x32_RUN16: (128, 32, 16), TILE_32x128_RUN16: (32, 128, 16),
               TILE_64x32_RUN8: (64, 32, 8)}


def _key(key):
    k = [int(v) for v in key] + [0] * (4 - len(key))
    return (ctypes.c_int64 * 4)(*k[:4])


def set_choice(knob: int, key, value: int) -> None:
    _lib.check(_lib.load().tg_tune_set(knob, _key(key), int(value)), "tg_tune_set")


def get_choice(knob: int, key) -> int | None:
    v = ctypes.c_int32(0)
    st = _lib.load().tg_tune_get(knob, _key(key), ctypes.byref(v))
    if st < 0:
        _lib.check(st, "tg_tune_get")
    return int(v.value) if st == 1 else None


def clear() -> None:
    _lib.load().tg_tune_clear()


def time_ms(fn, reps: int = 5, warmup: int = 1) -> float:
    """Median GPU time of fn() in ms (HIP events on the current stream)."""
    for _ in range(warmup):
        fn()
    times = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    return sorted(times)[len(times) // 2]


def _pick(knob, key, candidates, run, reps):
    """Time run() under each candidate choice; keep and return the fastest."""
    results = {}
    for c in candidates:
        set_choice(knob, key, c)
        results[int(c)] = time_ms(run, reps)
    best = min(results, key=results.get)
    set_choice(knob, key, best)
    return best, results


def tune_beamlet_sum(coef: torch.Tensor, n: int, shape, candidates=(8, 16, 32, 64, 128), reps: int = 3):
    """Beamlet slices for tg_beamlet_sum(n, H, W) on these coefficient records."""
    from .gaussian import BeamletSum

    summer = BeamletSum(n, shape, coef.device)
    cands = [c for c in candidates if c <= max(n, 1)] or [1]
    return _pick(BEAMLET_SPLITS, (n, shape[0], shape[1]), cands, lambda: summer.run(coef), reps)


def tune_beamlet_tile_shape(coef: torch.Tensor, n: int, shape, candidates=tuple(TILE_SHAPES), reps: int = 3):
    """Tile shape of tg_beamlet_sum(n, H, W)'s VALU recurrence on these
    coefficient records (matters only when that path runs)."""
    from .gaussian import BeamletSum

    summer = BeamletSum(n, shape, coef.device)
    return _pick(TILE_SHAPE, (n, shape[0], shape[1]), list(candidates), lambda: summer.run(coef), reps)


def tune_beamlet_path(coef: torch.Tensor, n: int, shape, candidates=(BEAMLET_PATH_AUTO, BEAMLET_PATH_VALU),
                      reps: int = 3):
    """Tile sum of tg_beamlet_sum(n, H, W) on these coefficient records: the
    matrix-core product (taken only when every beamlet is separable) or the
    VALU recurrence."""
    from .gaussian import BeamletSum

    summer = BeamletSum(n, shape, coef.device)
    return _pick(BEAMLET_PATH, (n, shape[0], shape[1]), list(candidates), lambda: summer.run(coef), reps)


def tune_run_to_image(plan, candidates=None, reps: int = 5):
    """Workgroups of a fused run_to_image plan (binning programs only)."""
    binning = [(prog, bs) for prog, bs, _ in plan.programs if bs is not None and bs.enabled]
    if not binning:
        raise ValueError("tune_run_to_image: the plan does not bin")
    prog, bs = binning[0]
    cus = int(_lib.load().tg_device_cu_count())
    if candidates is None:
        candidates = [cus // 2, cus, 2 * cus, 4 * cus]
    key = (prog.n, bs.height * bs.width, prog.dtype)
    return _pick(RAY_GRID, key, [c for c in candidates if c > 0], lambda: plan.launch(), reps)


def time_back_to_back_ms(fn, n: int) -> float:
    """GPU time per call of `n` back-to-back calls of fn() (one HIP event
    pair on the current stream around all of them)."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(n):
        fn()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / n


def tune_stem_fit(ws, params: torch.Tensor, targets: torch.Tensor, candidates=(STEM_FIT_SCAN, STEM_FIT_CAMERA),
                  steps: int = 64, rounds: int = 3):
    """Launch variant of tg_stem_fit for this geometry (ws: STEMFitWorkspace).

    Each candidate is timed as `steps` back-to-back fits, in `rounds`
    alternating rounds (best round kept): a fitting loop runs the fit back to
    back, and on sustained load the device settles into a different clock and
    power state than a single call after idle sees (the first ~30 ms of
    sustained load run up to ~25% slower on the descan fit).  The default
    budget (3 rounds x 2 candidates x 64 steps, ~70 ms) spans that settling."""
    g = ws.geometry
    key = (g.scan_h * g.scan_w, g.n_side * g.n_side, g.n_cam)
    results = {int(c): float("inf") for c in candidates}
    for _ in range(rounds):
        for c in candidates:
            set_choice(STEM_FIT, key, c)
            ws.run(params, targets)
            results[int(c)] = min(results[int(c)], time_back_to_back_ms(lambda: ws.run(params, targets), steps))
    best = min(results, key=results.get)
    set_choice(STEM_FIT, key, best)
    return best, results


def tune_gauss_fit(grids, counts, params: torch.Tensor, target: torch.Tensor, candidates=(128, 256, 512, 1024),
                   cutoff=None, reps: int = 3, shapes=tuple(TILE_SHAPES)):
    """Gaussians per forward slice, then the forward tile shape, for this
    batch geometry (each choice needs a new plan); returns
    (best slice, slice timings, batch planned with both best choices,
    best shape, shape timings)."""
    from .gauss_fit import GaussianFitBatch

    npix = sum(int(g.shape[0]) * int(g.shape[1]) for g in grids)
    key = (len(grids), int(sum(counts)), npix)
    shape_key = key + (1,)
    loss = torch.empty(len(grids), dtype=torch.float64, device=params.device)
    grad = torch.empty((int(sum(counts)), 16), dtype=torch.float64, device=params.device)

    def timed():
        b = GaussianFitBatch(grids, counts, cutoff)
        t = time_ms(lambda: b.run(params, target, None, loss, grad), reps)
        del b
        return t

    results = {}
    for c in candidates:
        set_choice(FIT_SLICE, key, c)
        results[int(c)] = timed()
    best = min(results, key=results.get)
    set_choice(FIT_SLICE, key, best)
    shape_ms = {}
    for sh in shapes:
        set_choice(TILE_SHAPE, shape_key, sh)
        shape_ms[int(sh)] = timed()
    best_shape = min(shape_ms, key=shape_ms.get)
    set_choice(TILE_SHAPE, shape_key, best_shape)
    return best, results, GaussianFitBatch(grids, counts, cutoff), best_shape, shape_ms


def save(path: str, entries) -> None:
    """Write [(knob, key, value), ...] as JSON (for reuse across processes)."""
    with open(path, "w") as f:
        json.dump([{"knob": _NAMES[k], "key": list(map(int, key)), "value": int(v)} for k, key, v in entries], f)


def load(path: str) -> int:
    """Record every entry of a file written by save(); returns the count."""
    inv = {v: k for k, v in _NAMES.items()}
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        set_choice(inv[r["knob"]], r["key"], r["value"])
    return len(rows)
