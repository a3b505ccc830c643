"""Benchmark of the batched propagation path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line (rank 0).  Workloads:

  stem_fit  (default) 4D-STEM descan/geometry fit, forward + adjoint: 64x64 scan
            positions x 4096 rays x 8 camera lengths (134,217,728 ray evaluations)
            against 1.07 GB of f32 target coordinates; loss + 26 gradients.
  stem_fwd  4D-STEM forward: 16.7M rays -> detector coords (f32) + bilinear
            sample values.
  plain     1,048,576 f32 rays through Lens(0.5, f=1) -> Detector(1.0, 128x128):
            fused run_to_image (positions, pathlength and int32 histogram).
  beamlet   Gaussian-beamlet aperture diffraction: 131,072 beamlets (300 keV,
            waist 50 nm, 10 um aperture) propagated 0.3 m onto 1024^2 pixels
            with the 4-sigma cut; complex64 field.
  gauss_fit batched Gaussian-decomposition fits, forward + adjoint: 256
            instances with 3-2,000 Gaussians (log-uniform, seed 5) on square
            detectors of 128-512 pixels (seed 6), 1 um pixels, 2.51 pm;
            per-instance intensity loss and the gradient of every parameter row.

A "step" is one pass of the hot path over one batch whose inputs are already
resident in HBM.  Timing: W untimed warm-up steps, then K steps bracketed by
torch.cuda.synchronize(); HIP events on the launch stream around every kernel
give its average duration for the roofline figure.  The CPU baseline is the
oracle timed on this host over a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
SKIP_CPU = False
BEAMLET_PATH = "tuned"  # --beamlet-path


def _profiled(fname, key, cast=float):
    """A `key: value` line of the PMC summary profiles/<fname> (rocprofv3
    cannot run inside this process); None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", fname)) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return cast(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def _profiled_traffic(fname):
    """HBM bytes per launch of a kernel as measured by the calibrated
    FETCH_SIZE + WRITE_SIZE PMC runs summarised in profiles/<fname>."""
    return _profiled(fname, "traffic_bytes_per_launch", int)


def _events(n):
    import torch

    return [torch.cuda.Event(enable_timing=True) for _ in range(n)]


def bench_plain(steps: int, warmup: int):
    import numpy as np
    import torch

    from oracle import ray_oracle as O
    from temgymcore_amd.components import Detector, Lens
    from temgymcore_amd.ray import Ray
    from temgymcore_amd.run import plan_run_to_image

    n = 1 << 20
    dev = torch.device("cuda", 0)
    x, y, dx, dy = O.plain_workload(n, seed=0)
    t = [torch.from_numpy(a).to(dev) for a in (x, y, dx, dy)]
    model = [Lens(z=0.5, focal_length=1.0), Detector(z=1.0, pixel_size=(0.01, 0.01), shape=(128, 128))]
    plan = plan_run_to_image(Ray(x=t[0], y=t[1], dx=t[2], dy=t[3], z=0.0, pathlength=0.0), model)
    from temgymcore_amd import autotune

    tuned_grid, tune_ms = autotune.tune_run_to_image(plan)

    def zero_and_launch():
        plan.acc.zero_()
        plan.launch()

    _settle(zero_and_launch)
    for _ in range(warmup):
        zero_and_launch()
    torch.cuda.synchronize()
    # kernel time: events around the launch alone
    ev = _events(2 * steps)
    for i in range(steps):
        plan.acc.zero_()
        ev[2 * i].record()
        plan.launch()
        ev[2 * i + 1].record()
    torch.cuda.synchronize()
    kern_ms = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(steps)) / steps

    # the timed steps: one step (zero the histogram + the fused launch) is
    # captured once in a HIP graph and replayed K times
    def step():
        plan.acc.zero_()
        plan.launch()

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    _settle(graph.replay)
    for _ in range(max(warmup, 1)):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        graph.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / steps
    bytes_per_launch = 36 * n + 128 * 128 * 4  # 16 B in + 20 B out per ray, + histogram
    achieved = bytes_per_launch / (kern_ms * 1e-3)
    # parity spot check of the timed buffers
    ref, hist = O.plain_run(x, y, dx, dy)
    parity = bool(np.array_equal(plan.out_ray.x.cpu().numpy(), ref["x"]) and
                  np.array_equal(plan.acc.cpu().numpy().astype(np.int64), np.asarray(hist, dtype=np.int64)))

    # CPU baseline: the numpy oracle on the full batch, one process per host core
    cpu = None
    if not SKIP_CPU:
        rate, cores = _cpu_parallel("plain", (n,))
        cpu = {"value": rate, "unit": "rays/s", "cores": cores, "kind": "port",
               "sample": f"numpy oracle, the full {n}-ray batch repeated for 8 s in each of {cores} "
                         f"single-threaded processes ({_cpu_info()[1]})"}
    return dict(
        metric="rays/s (plain run_to_image: 1,048,576 f32 rays, Lens + Detector 128x128 binning)",
        value=n / (ms_per_step * 1e-3), unit="rays/s", ms_per_step=ms_per_step, dtype="f32",
        config={"workload": "plain run_to_image", "rays": n, "detector": [128, 128], "seed": 0,
                "kernel_ms": kern_ms, "parity_x_and_histogram_bit_exact": parity,
                "timed_as": "K HIP-graph replays of (histogram zero + fused launch)",
                "autotune": {"ray_grid": tuned_grid, "candidate_ms": tune_ms}},
        roofline={"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK, "traffic": _profiled_traffic("ray_program_kernel-plain_1M.md"),
                  "traffic_source": "profiles/ray_program_kernel-plain_1M.md (calibrated FETCH_SIZE + WRITE_SIZE)"},
        cpu_baseline=cpu,
    )


def _cpu_pool_init():
    # one thread per worker process: the pool spreads the work over the cores
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = "1"


def _cpu_sample(kind, args):
    """(one repetition of a workload's oracle sample, work units per repetition)."""
    import numpy as np

    if kind == "plain":
        from oracle import ray_oracle as O

        (n,) = args
        x, y, dx, dy = O.plain_workload(n, seed=0)
        return (lambda: O.plain_run(x, y, dx, dy)), n
    if kind == "stem_fwd":
        from oracle import stem_oracle as SO

        geom, p, n_pos, n_rays = args
        idx = np.arange(n_pos)
        return (lambda: (SO.forward(geom, p, scan_idx=idx), SO.sample_plane(geom, p))), n_pos * n_rays
    if kind == "backproject":
        from oracle import stem_oracle as SO

        geom, p, D, dT, iT, units = args
        return (lambda: SO.backproject(geom, p, D, dT, iT)), units
    if kind == "beamlet":
        from oracle import beamlet_oracle as BO

        x, y, pix, H, W, L, lam = args
        nb = len(x)
        r2 = BO.grid_coords((pix, pix), (H, W))
        A = np.eye(5)
        A[0, 2] = A[1, 3] = L
        d = dict(dx=np.zeros(nb), dy=np.zeros(nb), pathlength=np.zeros(nb), amplitude=np.ones(nb),
                 waist_xy=np.full((nb, 2), 5e-8), radii=np.full((nb, 2), np.inf), wavelength=np.full(nb, lam),
                 theta=np.zeros(nb))
        return (lambda: BO.make_image(x, y, abcd=A, r2=r2, **d)), nb * H * W
    if kind == "gauss_fit":
        from oracle import beamlet_oracle as BO
        from oracle import gauss_fit_oracle as GF

        items = [(p, BO.grid_coords((1e-6, 1e-6), hw), t) for p, hw, t in args]

        def rep():
            for p, r1, t in items:
                GF.loss(p, r1, t)
                GF.grad(p, r1, t)

        return rep, sum(2 * len(p) * len(r1) for p, r1, _ in items)
    raise ValueError(kind)


def _cpu_sample_worker(job):
    """Repeat one oracle sample from `start_at` (wall clock) for `seconds`;
    returns (units done, seconds)."""
    kind, args, start_at, seconds = job
    rep, units = _cpu_sample(kind, args)
    rep()  # untimed warm-up
    while time.time() < start_at:
        time.sleep(0.001)
    done, t0 = 0, time.perf_counter()
    while True:
        rep()
        done += units
        el = time.perf_counter() - t0
        if el >= seconds:
            return done, el


def _cpu_parallel(kind, args, seconds=8.0):
    """The oracle sample `kind` on every host core: one single-threaded
    process per core, all starting together; returns (units/s, cores)."""
    import concurrent.futures as cf
    import multiprocessing as mp

    cores = len(os.sched_getaffinity(0))
    start_at = time.time() + 6.0  # after every worker has spawned and warmed up
    with cf.ProcessPoolExecutor(cores, mp_context=mp.get_context("spawn"), initializer=_cpu_pool_init) as ex:
        res = list(ex.map(_cpu_sample_worker, [(kind, args, start_at, seconds)] * cores))
    return sum(u / e for u, e in res), cores


def _cpu_info():
    cores = len(os.sched_getaffinity(0))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def hbm_read_probe(n_bytes: int = 2 << 30, reps: int = 20):
    """Achievable HBM read rate (B/s) of this device, measured with the
    library's streaming probe kernel over a 2 GiB buffer."""
    import ctypes

    import torch

    from temgymcore_amd import _lib

    lib = _lib.load()
    dev = torch.device("cuda", 0)
    buf = torch.ones(n_bytes // 4, dtype=torch.float32, device=dev)
    n_blocks = 64 * lib.tg_device_cu_count()
    sums = torch.empty(n_blocks, dtype=torch.float32, device=dev)
    stream = _lib.stream_ptr(dev)

    def run():
        _lib.check(lib.tg_hbm_read_probe(buf.data_ptr(), n_bytes, sums.data_ptr(), n_blocks, stream),
                   "tg_hbm_read_probe")

    for _ in range(3):
        run()
    torch.cuda.synchronize()
    # one event pair around `reps` back-to-back launches (per-launch pairs
    # add the launch gap to every ~0.3 ms measurement)
    e0, e1 = _events(2)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    ok = abs(float(sums.double().sum()) - n_bytes / 4) < 1e-6 * n_bytes
    del buf
    return n_bytes / (ms * 1e-3), ok


SETTLE_MS = 60.0


def _settle(step_fn, ms=SETTLE_MS):
    """Untimed back-to-back steps for about `ms` of GPU time.  After idle the
    device's power management takes ~30 ms of sustained load to settle, and
    meanwhile an HBM-bound pass runs up to ~25% slower (measured on the descan
    fit: 20-launch medians of 181, 214, 192, 180, 174 ... 165 us after 0.5 s
    idle).  A fitting loop runs in the settled state, which is what the
    timed steps measure."""
    import torch

    e0, e1 = _events(2)
    e0.record()
    step_fn()
    e1.record()
    e1.synchronize()
    n = int(min(5000, max(1, ms / max(e0.elapsed_time(e1), 1e-3))))
    for _ in range(n):
        step_fn()
    return n


def _time_steps(step_fn, steps, warmup, graph=True, settle=True):
    """Time `steps` back-to-back steps on the current stream.

    Returns (wall ms/step, GPU ms/step from one HIP event pair around the K
    steps, GPU ms/step of the same K steps as HIP-graph replays, or None).
    With `settle`, ~60 ms of untimed steps (_settle) run before the W warmup
    steps and before the graph replays."""
    import torch

    import gc

    e0, e1 = _events(2)
    # no gc.collect() here: a collection right before the timed region leaves
    # the GPU idle long enough that the timed steps start at lowered clocks
    # (measured: 180 vs 165 us per fit pass); disabling the collector keeps
    # its pauses out of the warmup and the timed region instead
    gc.disable()
    try:
        if settle:
            _settle(step_fn)
        for _ in range(max(warmup, 1)):
            step_fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0.record()
        for _ in range(steps):
            step_fn()
        e1.record()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / steps
    finally:
        gc.enable()
    gpu = e0.elapsed_time(e1) / steps
    if not graph:
        return wall, gpu, None
    # the same step captured once in a HIP graph and replayed K times
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        step_fn()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step_fn()
    if settle:
        _settle(graph.replay)
    for _ in range(max(warmup, 1)):
        graph.replay()
    torch.cuda.synchronize()
    g0, g1 = _events(2)
    g0.record()
    for _ in range(steps):
        graph.replay()
    g1.record()
    torch.cuda.synchronize()
    return wall, gpu, g0.elapsed_time(g1) / steps


def bench_stem_fit(steps: int, warmup: int):
    import ctypes

    import numpy as np
    import torch

    from oracle import c as cport
    from temgymcore_amd import stem

    dev = torch.device("cuda", 0)
    model, _ = stem.reference_workload(n_cam=8)
    params = stem.params_tensor(model, dev)
    targets = stem.stem_forward(stem.perturbed(model))["coords"]
    ws = stem.STEMFitWorkspace(model, dev)
    units = model.n_cam * model.n_scan * model.n_rays
    from temgymcore_amd import autotune

    tuned_variant, tune_ms = autotune.tune_stem_fit(ws, params, targets)
    ms_per_step, step_ms, graph_ms = _time_steps(lambda: ws.run(params, targets), steps, warmup)
    gpu_out = ws.run(params, targets).cpu().numpy()
    # the dominant kernel alone: the same K launches of the per-scan-position
    # pass (tg_stem_fit_terms), bracketed by one HIP event pair on its stream
    _, kern_ms, _ = _time_steps(lambda: ws.terms(params, targets), steps, warmup, graph=False)
    bytes_per_launch = 8 * units  # f32 (x, y) target per ray and camera length, read once
    achieved = bytes_per_launch / (kern_ms * 1e-3)
    probe_bw, probe_ok = hbm_read_probe()

    names = {1: "scan", 2: "camera"}
    tune_info = {"stem_fit_variant": names[tuned_variant], "candidate_ms": {names[k]: v for k, v in tune_ms.items()},
                 "timed_as": "64 back-to-back steps per candidate, best of 3 alternating rounds"}
    oracle_err = _stem_fit_oracle_error(model, ws, targets) if not SKIP_CPU else None
    if SKIP_CPU:
        res = _result_fit(units, ms_per_step, step_ms, kern_ms, graph_ms, gpu_out, achieved, probe_bw, probe_ok, None,
                          None)
        res["config"]["autotune"] = tune_info
        return res
    # CPU baseline: the OpenMP C port on a bounded sample of scan positions
    lib = cport.load()
    cores, cpu_model = _cpu_info()
    t_host = targets.cpu().numpy()
    off = np.zeros(4)
    p_host = model.params()
    out = np.zeros(ws.np)
    n_s, elapsed, reps = 32, 0.0, 0
    while elapsed < 10.0:
        c0 = time.perf_counter()
        lib.stem_fit(64, 64, 64, 8, model.semi_conv, model.defocus, model.pivot, off.ctypes.data,
                     p_host.ctypes.data, t_host.ctypes.data, 0, n_s, out.ctypes.data)
        elapsed += time.perf_counter() - c0
        reps += 1
        if elapsed < 1.0 and n_s < model.n_scan:
            n_s, elapsed, reps = min(model.n_scan, n_s * 4), 0.0, 0
    cpu_rate = reps * n_s * model.n_rays * model.n_cam / elapsed
    # parity of the timed run on the same sample: GPU partial rows vs the CPU port
    rows = ws.partials.cpu().numpy()[:, :n_s].sum(axis=1)
    rel = float(np.max(np.abs(rows - out) / np.maximum(np.abs(out), 1e-300)))
    cpu = {"value": cpu_rate, "unit": "ray-evals/s", "cores": cores, "kind": "port",
           "sample": f"OpenMP C port, {n_s} scan positions x 4096 rays x 8 camera lengths, "
                     f"{reps} reps ({cpu_model})"}
    res = _result_fit(units, ms_per_step, step_ms, kern_ms, graph_ms, gpu_out, achieved, probe_bw, probe_ok, cpu, rel)
    res["config"]["autotune"] = tune_info
    res["config"]["oracle_error"] = oracle_err
    return res


def _stem_fit_oracle_error(model, ws, targets, n_pick: int = 4):
    """Error of the timed result against the float64 oracle, outside the timed
    region: the GPU's per-scan-position rows (loss + 26 gradient entries, from
    the timed buffers) summed over `n_pick` seeded scan positions, against the
    oracle's loss and its central-difference gradient (one Richardson step) on
    the same f32 targets."""
    import numpy as np

    from oracle import stem_oracle as SO

    geom = dict(scan_shape=model.scan_shape, n_side=model.n_side, semi_conv=model.semi_conv, defocus=model.defocus,
                pivot=model.pivot, offsets=tuple(model.descan_offsets))
    pick = np.sort(np.random.default_rng(0).choice(model.n_scan, n_pick, replace=False))
    rows = ws.partials.cpu().numpy().T[pick].sum(axis=0)  # [np]: 26 gradient entries, then the loss
    t = targets.cpu().numpy()[:, pick]
    ref_loss = SO.loss(geom, model.params(), t, scan_idx=pick)
    ref_g = SO.fd_grad(geom, model.params(), t, scan_idx=pick)
    blocks = {"descan_16": slice(0, 16), "scan_step": slice(16, 17), "scan_rotation": slice(17, 18),
              "camera_lengths_8": slice(18, 18 + model.n_cam)}
    grad = {k: float(np.max(np.abs(rows[b] - ref_g[b])) / np.max(np.abs(ref_g[b]))) for k, b in blocks.items()}
    return {"scan_positions": [int(s) for s in pick], "loss_rel": float(abs(rows[-1] - ref_loss) / abs(ref_loss)),
            "grad_max_abs_err_over_block_max": grad,
            "oracle": "float64 numpy restatement; gradient by central differences with one Richardson step"}


def _result_fit(units, ms_per_step, step_ms, kern_ms, graph_ms, gpu_out, achieved, probe_bw, probe_ok, cpu, rel):
    step_bw = 8 * units / (ms_per_step * 1e-3)
    return dict(
        metric="rays/s fwd+bwd (16.7M-ray 4D-STEM model x 8 camera lengths: descan/geometry fit loss + 26 grads)",
        value=units / (ms_per_step * 1e-3), unit="ray-evals/s", ms_per_step=ms_per_step, dtype="f32 (f64 reductions)",
        config={"workload": "stem_fit: 64x64 scan, 64x64 rays, 8 camera lengths, 1.07 GB f32 targets",
                "units_per_step": units, "event_ms_per_step": step_ms, "graph_replay_ms_per_step": graph_ms,
                "dominant_kernel_event_ms": kern_ms,
                "kernels_per_step": "stem_fit_kernel + row_sum_kernel, or stem_fit_cam_kernel + stem_fit_epilogue + "
                                    "row_sum_kernel (the tuned variant); roofline.achieved uses the event time of "
                                    "the per-scan-position pass alone (tg_stem_fit_terms)",
                "cpu_port_rel_diff_on_sample": rel, "loss": float(gpu_out[-1])},
        roofline={"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK,
                  "traffic": _profiled_traffic("stem_fit_kernel-stem_fit_64x64x4096x8.md"),
                  "traffic_source": "profiles/stem_fit_kernel-stem_fit_64x64x4096x8.md (calibrated FETCH_SIZE + WRITE_SIZE)",
                  "measured_read_peak": probe_bw / 1e9 if probe_ok else None,
                  "frac_of_measured": achieved / probe_bw if probe_ok else None,
                  "step_achieved": step_bw / 1e9, "step_frac": step_bw / HBM_PEAK,
                  "step_note": "the whole timed step (fit pass + final column sums + launch gaps): algorithmic "
                               "bytes / ms_per_step"},
        cpu_baseline=cpu,
    )


def bench_stem_fwd(steps: int, warmup: int):
    import numpy as np
    import torch

    from oracle import stem_oracle as SO
    from temgymcore_amd import stem

    dev = torch.device("cuda", 0)
    model, sample = stem.reference_workload(n_cam=1)
    params = stem.params_tensor(model, dev)
    sample_t = torch.from_numpy(sample).to(dev)
    out = stem.stem_forward(model, params, sample_t, coords=True, values=True)
    units = model.n_scan * model.n_rays
    ms_per_step, kern_ms, graph_ms = _time_steps(
        lambda: stem.stem_forward(model, params, sample_t, coords=True, values=True, out=out), steps, warmup)
    bytes_per_launch = 12 * units + sample.nbytes  # (x, y, value) f32 written per ray + image read once
    achieved = bytes_per_launch / (kern_ms * 1e-3)
    geom = dict(scan_shape=model.scan_shape, n_side=model.n_side, semi_conv=model.semi_conv,
                defocus=model.defocus, pivot=model.pivot, offsets=model.descan_offsets)
    cpu = None
    if not SKIP_CPU:
        rate, cores = _cpu_parallel("stem_fwd", (geom, model.params(), 16, model.n_rays))
        cpu = {"value": rate, "unit": "rays/s", "cores": cores, "kind": "port",
               "sample": f"numpy f64 oracle, 16 scan positions x 4096 rays (+ the sample-plane lookup) repeated "
                         f"for 8 s in each of {cores} single-threaded processes ({_cpu_info()[1]})"}
    return dict(
        metric="rays/s (4D-STEM shadow-image forward: 16.7M rays -> detector coords + bilinear sample values)",
        value=units / (ms_per_step * 1e-3), unit="rays/s", ms_per_step=ms_per_step, dtype="f32",
        config={"workload": "stem_fwd: 64x64 scan, 64x64 rays, camera length 0.5 m", "rays": units,
                "event_ms_per_step": kern_ms, "graph_replay_ms_per_step": graph_ms},
        roofline={"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK,
                  "traffic": _profiled_traffic("stem_forward_kernel-stem_fwd_64x64x4096.md"),
                  "traffic_source": "profiles/stem_forward_kernel-stem_fwd_64x64x4096.md (calibrated FETCH_SIZE + "
                                    "WRITE_SIZE)"},
        cpu_baseline=cpu,
    )


def bench_backproject(steps: int, warmup: int):
    import numpy as np
    import torch

    from oracle import ray_oracle as R
    from oracle import stem_oracle as SO
    from temgymcore_amd import stem

    dev = torch.device("cuda", 0)
    model, sample = stem.reference_workload(n_cam=1)
    params = stem.params_tensor(model, dev)
    data = stem.stem_forward(model, params, torch.from_numpy(sample).to(dev), coords=False,
                             image4d=True)["image4d"]  # 1 GiB f32 shadow images
    units = model.n_scan * model.n_rays
    img = torch.zeros(tuple(model.sample_shape), dtype=torch.float64, device=dev)
    g = model.geometry()
    lib = stem._lib.load()

    def step():
        img.zero_()
        st = lib.tg_stem_backproject(stem.ctypes.byref(g), params.data_ptr(), data.data_ptr(), 0, img.data_ptr(),
                                     None, stem._lib.stream_ptr(dev))
        stem._lib.check(st, "tg_stem_backproject")

    ms_per_step, kern_ms, graph_ms = _time_steps(step, steps, warmup)
    # algorithmic bytes: the one f32 detector value each ray reads
    bytes_per_launch = 4 * units
    achieved = bytes_per_launch / (kern_ms * 1e-3)
    cpu = None
    if not SKIP_CPU:
        small, _ = stem.reference_workload(n_cam=1, scan_shape=(4, 4))
        geom = dict(scan_shape=small.scan_shape, n_side=small.n_side, semi_conv=small.semi_conv,
                    defocus=small.defocus, pivot=small.pivot, offsets=small.descan_offsets,
                    sample_shape=tuple(small.sample_shape))
        D = data[:small.n_scan].double().cpu().numpy()
        dT = R.metres_to_pixels_matrix((0.0, 0.0), (small.det_pixel,) * 2, small.det_shape)
        iT = R.metres_to_pixels_matrix((0.0, 0.0), (small.sample_pixel,) * 2, small.sample_shape)
        rate, cores = _cpu_parallel("backproject", (geom, small.params(), D, dT, iT, small.n_scan * small.n_rays))
        cpu = {"value": rate, "unit": "rays/s", "cores": cores, "kind": "port",
               "sample": f"numpy f64 oracle, 16 scan positions x 4096 rays repeated for 8 s in each of {cores} "
                         f"single-threaded processes ({_cpu_info()[1]})"}
    return dict(
        metric="rays/s (4D-STEM shadow-image back-projection: 16.7M rays of a 1 GiB dataset onto the sample plane)",
        value=units / (ms_per_step * 1e-3), unit="rays/s", ms_per_step=ms_per_step, dtype="f32 data, f64 sums",
        config={"workload": "backproject: 64x64 scan, 64x64 rays, 256x256 detector, 512x512 sample",
                "rays": units, "event_ms_per_step": kern_ms, "graph_replay_ms_per_step": graph_ms,
                "step": "zero the f64 image + one back-projection launch"},
        roofline={"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK,
                  "traffic": _profiled_traffic("stem_backproject_kernel-backproject_64x64x4096.md"),
                  "traffic_source": "profiles/stem_backproject_kernel-backproject_64x64x4096.md (calibrated "
                                    "FETCH_SIZE; 4x the algorithmic bytes: one cache line per gathered value)",
                  "note": "gather of one f32 per ray; adjacent rays land four detector pixels apart, so "
                          "each 64 B sector serves four rays (traffic 4x algorithmic, ~4.7 TB/s of sector "
                          "reads), co-limited by the per-ray FP64 position and 3x3 window arithmetic"},
        cpu_baseline=cpu,
    )


def aperture_beamlets(n=131072, seed=4):
    """Seeded beamlet centres uniform in a 10 um diameter disc (radius 5 um)."""
    import numpy as np

    rng = np.random.default_rng(seed)
    r = 5e-6 * np.sqrt(rng.uniform(0, 1, n))
    a = rng.uniform(0, 2 * np.pi, n)
    return r * np.cos(a), r * np.sin(a)


def _ellipse_pixel_count(coef_bytes, n, H, W):
    """Exact number of (beamlet, pixel) pairs inside the cutoff ellipses."""
    import numpy as np

    rec = np.frombuffer(coef_bytes, dtype=np.dtype([("phi", "<f8", 12), ("k", "<f4", 2), ("c", "<f4", 2),
                                                    ("abg", "<f4", 3), ("cut", "<f4"), ("box", "<f4", 2)]))
    cx, cy = rec["c"][:, 0].astype(np.float64), rec["c"][:, 1].astype(np.float64)
    al, be, ga = (rec["abg"][:, i].astype(np.float64) for i in range(3))
    total = 0
    rows = np.arange(H, dtype=np.float64)
    for lo in range(0, n, 4096):
        sl = slice(lo, min(n, lo + 4096))
        v = rows[None, :] - cy[sl, None]
        t2 = be[sl, None] - ga[sl, None] * v * v
        ok = t2 >= 0
        sq = np.sqrt(np.where(ok, t2, 0.0))
        mid = cx[sl, None] - al[sl, None] * v
        lo_px = np.clip(np.ceil(mid - sq), 0, W)
        hi_px = np.clip(np.floor(mid + sq), -1, W - 1)
        total += int(np.where(ok, np.maximum(hi_px - lo_px + 1, 0), 0).sum())
    return total


def _shape_name(shape):
    from temgymcore_amd import autotune

    w, h, run = autotune.TILE_SHAPES[int(shape)]
    return f"{w}x{h} run {run}"


def bench_beamlet(steps: int, warmup: int):
    import numpy as np
    import torch

    from oracle import beamlet_oracle as BO
    from temgymcore_amd.components import Detector
    from temgymcore_amd.gaussian import (MODE_PROPAGATED, BeamletSum, GaussianRay, _abcd_per_beamlet,
                                         beamlet_coefficients, pack_params)

    dev = torch.device("cuda", 0)
    n, H, W, pix, L, lam = 131072, 1024, 1024, 4e-8, 0.3, 1.97e-12
    x, y = aperture_beamlets(n)
    z = np.zeros(n)
    rays = GaussianRay(x=torch.from_numpy(x), y=torch.from_numpy(y), dx=torch.from_numpy(z), dy=torch.from_numpy(z),
                       z=torch.from_numpy(z), pathlength=torch.from_numpy(z), _one=torch.ones(n, dtype=torch.float64),
                       amplitude=torch.ones(n, dtype=torch.float64), waist_xy=torch.full((n, 2), 5e-8, dtype=torch.float64),
                       radii_of_curv=torch.full((n, 2), float("inf"), dtype=torch.float64),
                       wavelength=torch.full((n,), lam, dtype=torch.float64), theta=torch.from_numpy(z))
    det = Detector(z=L, pixel_size=(pix, pix), shape=(H, W))
    params = pack_params(rays, dev)
    abcd, stride = _abcd_per_beamlet(rays, [det], dev, n)
    from temgymcore_amd import autotune

    coef0 = beamlet_coefficients(params, det, MODE_PROPAGATED, abcd, stride, cutoff=4.0)
    # the tile-sum path (matrix cores for separable beamlets, or the VALU
    # recurrence), then the beamlet slices under the chosen path
    if BEAMLET_PATH == "tuned":
        tuned_path, path_ms = autotune.tune_beamlet_path(coef0, n, (H, W))
    else:
        tuned_path = autotune.BEAMLET_PATH_AUTO if BEAMLET_PATH == "auto" else autotune.BEAMLET_PATH_VALU
        autotune.set_choice(autotune.BEAMLET_PATH, (n, H, W), tuned_path)
        path_ms = {}
    tuned_splits, tune_ms = autotune.tune_beamlet_sum(coef0, n, (H, W))
    tuned_shape, shape_ms = None, {}
    if tuned_path == autotune.BEAMLET_PATH_VALU:
        # the recurrence's tile shape (the matrix-core path has one shape)
        tuned_shape, shape_ms = autotune.tune_beamlet_tile_shape(coef0, n, (H, W))
    summer = BeamletSum(n, (H, W), dev)
    holder = {}

    def step():
        coef = beamlet_coefficients(params, det, MODE_PROPAGATED, abcd, stride, cutoff=4.0)
        holder["coef"] = coef
        summer.run(coef)

    ms_per_step, kern_ms, graph_ms = _time_steps(step, steps, warmup)
    # tg_beamlet_sum alone (path flags, tile sum, slice sum) on the same records
    _, sum_ms, _ = _time_steps(lambda: summer.run(coef0), steps, warmup, graph=False)
    rec_bytes = holder["coef"].cpu().numpy().tobytes()
    evaluated = _ellipse_pixel_count(rec_bytes, n, H, W)
    dense = n * H * W
    evals_per_s = evaluated / (ms_per_step * 1e-3)
    matrix = tuned_path == autotune.BEAMLET_PATH_AUTO and _all_separable(rec_bytes)
    if matrix:
        # the matrix-core product computes whole 64x64 tiles: 8 flops (one
        # complex multiply-add) per (beamlet, pixel) of every tile its cutoff box meets
        tile_pairs = _tile_pair_count(rec_bytes, n, H, W)
        achieved = 8 * tile_pairs / (sum_ms * 1e-3)
        roof = {"bound": "mfma", "achieved": achieved / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                "frac": achieved / 157.3e12, "traffic": None, "flops_per_pair": 8, "tile_pairs": tile_pairs,
                "timed_kernels": "tg_beamlet_sum (path flags + beamlet_sum_sep_kernel + slice sums), HIP events",
                "peak_note": "MI355X dense FP32 matrix peak (v_mfma_f32_32x32x2f32)",
                "mfma_util_percent": _profiled("beamlet_sum_sep_kernel-beamlet_131072_1024sq.md",
                                               "mfma_util_percent"),
                "mfma_source": "profiles/beamlet_sum_sep_kernel-beamlet_131072_1024sq.md (MfmaUtil PMC); the "
                               "factor evaluation's vector instructions do not overlap the f32 MFMAs on gfx950"}
    else:
        flops_per_eval = 14  # two complex f32 products (6 each) + complex accumulate (2)
        achieved = flops_per_eval * evaluated / (sum_ms * 1e-3)
        roof = {"bound": "valu", "achieved": achieved / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                "frac": achieved / 157.3e12, "traffic": None, "flops_per_eval": flops_per_eval,
                "valu_busy_percent": _profiled("beamlet_sum_kernel-beamlet_131072_1024sq.md", "valu_busy_percent"),
                "valu_source": "profiles/beamlet_sum_kernel-beamlet_131072_1024sq.md (VALUBusy PMC): the sum is "
                               "bound by VALU issue; frac counts only the useful FP32 flops"}
    cpu = None
    if not SKIP_CPU:
        nb = 16
        rate, cores = _cpu_parallel("beamlet", (x[:nb], y[:nb], pix, H, W, L, lam))
        cpu = {"value": rate, "unit": "beamlet-pixel evals/s (dense)", "cores": cores, "kind": "port",
               "sample": f"numpy complex128 oracle (reference formulas), {nb} beamlets x the full 1024^2 detector "
                         f"repeated for 8 s in each of {cores} single-threaded processes ({_cpu_info()[1]})"}
    return dict(
        metric="beamlet-pixel evals/s (131,072-beamlet aperture diffraction, 1024^2 detector, 4-sigma cut)",
        value=evals_per_s, unit="beamlet-pixel evals/s", ms_per_step=ms_per_step, dtype="f32 (f64 phase)",
        config={"workload": "beamlet: 131072 beamlets, 10 um aperture, w0 50 nm, 0.3 m, 1024^2 px of 40 nm",
                "evaluated_pairs": evaluated, "dense_pairs": dense, "dense_equiv_evals_per_s": dense / (ms_per_step * 1e-3),
                "event_ms_per_step": kern_ms, "graph_replay_ms_per_step": graph_ms, "beamlet_sum_event_ms": sum_ms,
                "tile_sum": "matrix cores (separable beamlets)" if matrix else "VALU recurrence",
                "kernels_per_step": "beamlet_coef_kernel + sep_flag_kernel + beamlet_sum_sep_kernel / "
                                    "beamlet_sum_kernel (one exits at once) + split_sum_kernel",
                "autotune": {"beamlet_path": {1: "auto", 2: "valu"}[tuned_path],
                             "path_candidate_ms": {{1: "auto", 2: "valu"}[k]: v for k, v in path_ms.items()},
                             "beamlet_splits": tuned_splits, "candidate_ms": tune_ms,
                             "tile_shape": _shape_name(tuned_shape) if tuned_shape else "64x64 (matrix cores)",
                             "tile_shape_ms": {_shape_name(k): v for k, v in shape_ms.items()}}},
        roofline=roof,
        cpu_baseline=cpu,
    )


_COEF_DTYPE = [("phi", "<f8", 12), ("k", "<f4", 2), ("c", "<f4", 2), ("abg", "<f4", 3), ("cut", "<f4"),
               ("box", "<f4", 2)]


def _all_separable(coef_bytes):
    """Every record separable as the tile-sum path test has it (finite Phi,
    no px*py term, decaying along both axes)."""
    import numpy as np

    phi = np.frombuffer(coef_bytes, dtype=np.dtype(_COEF_DTYPE))["phi"]
    return bool(np.all(np.isfinite(phi)) and np.all(phi[:, 2] == 0) and np.all(phi[:, 3] == 0)
                and np.all(phi[:, 0] < 0) and np.all(phi[:, 4] < 0))


def _tile_pair_count(coef_bytes, n, H, W, tile=64):
    """(beamlet, pixel) pairs of the 64x64 tiles each separable beamlet's
    cutoff ellipse meets (the matrix-core tile sum's list test: the tile's
    pixel centre nearest the centre inside u^2 + gamma v^2 <= beta, in f32),
    4096 per tile."""
    import numpy as np

    rec = np.frombuffer(coef_bytes, dtype=np.dtype(_COEF_DTYPE))
    f32 = np.float32
    cx, cy = rec["c"][:, 0].astype(f32), rec["c"][:, 1].astype(f32)
    gamma, beta = rec["abg"][:, 2].astype(f32), rec["abg"][:, 1].astype(f32)
    tx0 = (np.arange(-(-W // tile)) * tile).astype(f32)
    ty0 = (np.arange(-(-H // tile)) * tile).astype(f32)
    total = 0
    for lo in range(0, n, 4096):
        sl = slice(lo, min(n, lo + 4096))
        du = np.maximum(np.maximum(tx0[None, :] - cx[sl, None], cx[sl, None] - (tx0[None, :] + f32(tile - 1))), f32(0))
        dv = np.maximum(np.maximum(ty0[None, :] - cy[sl, None], cy[sl, None] - (ty0[None, :] + f32(tile - 1))), f32(0))
        gv = (gamma[sl, None] * dv).astype(np.float64) * dv  # fma(gamma dv, dv, du^2) in one rounding
        q = (gv[:, :, None] + (du * du).astype(np.float64)[:, None, :]).astype(f32)
        total += int((q <= beta[sl, None, None]).sum())
    return total * tile * tile


def bench_gauss_fit(steps: int, warmup: int):
    import numpy as np
    import torch

    from oracle import beamlet_oracle as BO
    from oracle import gauss_fit_oracle as GF
    from temgymcore_amd import gauss_fit

    from temgymcore_amd import autotune

    batch, params, target = gauss_fit.workload()
    tuned_slice, tune_ms, batch, tuned_shape, shape_ms = autotune.tune_gauss_fit(batch.grids, batch.counts, params,
                                                                                 target)
    dev = batch.device
    loss = torch.empty(batch.n_inst, dtype=torch.float64, device=dev)
    grad = torch.empty((batch.n_gauss, 16), dtype=torch.float64, device=dev)

    def step():
        batch.run(params, target, None, loss, grad)

    def fwd():
        batch.run(params, target, None, loss, None)

    ms_per_step, kern_ms, graph_ms = _time_steps(step, steps, warmup)
    fwd_ms, fwd_kern_ms, _ = _time_steps(fwd, max(steps // 2, 1), 1)
    torch.cuda.synchronize()
    assert torch.isfinite(loss).all() and torch.isfinite(grad).all()
    # pixels inside each Gaussian's cutoff ellipse, clipped to its detector
    rec = batch.workspace[batch.plan.off_coef:batch.plan.off_coef + 136 * batch.n_gauss].cpu().numpy().tobytes()
    evaluated = 0
    for i in range(batch.n_inst):
        g0, g1 = int(batch.gauss_offsets[i]), int(batch.gauss_offsets[i + 1])
        if g1 > g0:
            evaluated += _ellipse_pixel_count(rec[136 * g0:136 * g1], g1 - g0, *batch.shapes[i])
    dense = int(sum(c * h * w for c, (h, w) in zip(batch.counts, batch.shapes)))
    units = 2 * dense  # forward + backward Gaussian-pixel evaluations, dense count
    value = units / (ms_per_step * 1e-3)
    flops_fwd, flops_bwd = 14, 28  # per evaluated pair (recurrence + accumulate; + W product and 3 moments)
    achieved = (flops_fwd + flops_bwd) * evaluated / (kern_ms * 1e-3)
    cpu = None
    if not SKIP_CPU:
        # the oracle's loss + gradient over the first (up to) 16 Gaussians of
        # the first 6 instances, each on its full detector
        p = params.cpu().numpy()
        t = target.cpu().numpy().astype(np.float64)
        items, names = [], []
        for i in range(6):
            g0 = int(batch.gauss_offsets[i])
            g1 = min(int(batch.gauss_offsets[i + 1]), g0 + 16)
            h, w = batch.shapes[i]
            items.append((p[g0:g1], (h, w), t[int(batch.pix_offsets[i]):int(batch.pix_offsets[i + 1])]))
            names.append(f"{g1 - g0}x{h}x{w}")
        rate, cores = _cpu_parallel("gauss_fit", items)
        cpu = {"value": rate, "unit": "Gaussian-pixel evals/s fwd+bwd (dense)", "cores": cores, "kind": "port",
               "sample": f"numpy complex128 oracle loss + analytic gradient, the first <=16 Gaussians of instances "
                         f"0..5 ({', '.join(names)}), repeated for 8 s in each of {cores} single-threaded processes "
                         f"({_cpu_info()[1]})"}
    return dict(
        metric="Gaussian-pixel evals/s fwd+bwd (256 batched Gaussian-decomposition fits)",
        value=value, unit="Gaussian-pixel evals/s", ms_per_step=ms_per_step, dtype="f32 (f64 phase, reductions)",
        config={"workload": "gauss_fit: 256 instances, 3-2000 Gaussians (log-uniform, seed 5), "
                            "128-512 px square detectors (seed 6), 1 um pixels, 2.51 pm, no cutoff",
                "gaussians": batch.n_gauss, "pixels": batch.n_pix, "dense_pairs": dense,
                "evaluated_pairs_per_pass": evaluated, "event_ms_per_step": kern_ms,
                "graph_replay_ms_per_step": graph_ms, "forward_only_event_ms": fwd_kern_ms,
                "kernels_per_step": "fit_coef, fit_forward, fit_loss, fit_loss_final, fit_moment, fit_grad",
                "autotune": {"fit_slice": tuned_slice, "candidate_ms": tune_ms,
                             "tile_shape": _shape_name(tuned_shape),
                             "tile_shape_ms": {_shape_name(k): v for k, v in shape_ms.items()}}},
        roofline={"bound": "valu", "achieved": achieved / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                  "frac": achieved / 157.3e12, "traffic": None,
                  "flops_per_eval": {"forward": flops_fwd, "backward": flops_bwd},
                  "kernels": "fit_forward_kernel + fit_moment_kernel (whole step timed)",
                  "valu_busy_percent": _profiled("fit_forward_moment_kernels-gauss_fit_256.md", "valu_busy_percent"),
                  "valu_source": "profiles/fit_forward_moment_kernels-gauss_fit_256.md (VALUBusy PMC of the "
                                 "forward kernel)"},
        cpu_baseline=cpu,
    )


WORKLOADS = {"stem_fit": bench_stem_fit, "stem_fwd": bench_stem_fwd, "backproject": bench_backproject, "plain": bench_plain, "beamlet": bench_beamlet,
             "gauss_fit": bench_gauss_fit}
DEFAULT = "stem_fit"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU leg (for profiler runs; the JSON line then has cpu_baseline null)")
    ap.add_argument("--beamlet-path", default="tuned", choices=["tuned", "auto", "valu"],
                    help="beamlet workload: the tile-sum path (tuned: timed choice; auto: matrix cores for "
                         "separable beamlets; valu: the recurrence kernel)")
    args = ap.parse_args()
    global SKIP_CPU, BEAMLET_PATH
    SKIP_CPU = args.no_cpu_baseline
    BEAMLET_PATH = args.beamlet_path
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path would shard by scan position / detector tile with no exchange)",
              file=sys.stderr)
        sys.exit(2)
    res = WORKLOADS[args.workload](args.steps, args.warmup)
    res["config"]["untimed_settle"] = (f"~{SETTLE_MS:.0f} ms of back-to-back untimed steps before the W warmup steps "
                                       "of every timed measurement: the device's power management settles over "
                                       "~30 ms of sustained load after idle")
    line = {
        "metric": res["metric"], "value": res["value"], "unit": res["unit"], "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": res["dtype"],
        "data": "synthetic (seeded numpy default_rng)", "config": res["config"],
        "roofline": res["roofline"], "cpu_baseline": res["cpu_baseline"],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
