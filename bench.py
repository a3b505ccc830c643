"""Benchmark of the batched propagation path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Prints ONE JSON line (rank 0).  Workloads:

  plain  1,048,576 f32 rays through Lens(0.5, f=1) -> Detector(1.0, 128x128):
         fused run_to_image (positions, pathlength and int32 histogram).

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
    for _ in range(warmup):
        plan.acc.zero_()
        plan.launch()
    torch.cuda.synchronize()
    ev = _events(2 * steps)
    t0 = time.perf_counter()
    for i in range(steps):
        plan.acc.zero_()
        ev[2 * i].record()
        plan.launch()
        ev[2 * i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(steps)) / steps
    ms_per_step = wall * 1e3 / steps
    bytes_per_launch = 36 * n + 128 * 128 * 4  # 16 B in + 20 B out per ray, + histogram
    achieved = bytes_per_launch / (kern_ms * 1e-3)
    # parity spot check of the timed buffers
    ref, hist = O.plain_run(x, y, dx, dy)
    parity = bool(np.array_equal(plan.out_ray.x.cpu().numpy(), ref["x"]))

    # CPU baseline: the numpy oracle (single thread) on the full batch, ~10 s
    reps, c0 = 0, time.perf_counter()
    while time.perf_counter() - c0 < 10.0:
        O.plain_run(x, y, dx, dy)
        reps += 1
    cpu_s = (time.perf_counter() - c0) / reps
    return dict(
        metric="rays/s (plain run_to_image: 1,048,576 f32 rays, Lens + Detector 128x128 binning)",
        value=n / (ms_per_step * 1e-3), unit="rays/s", ms_per_step=ms_per_step, dtype="f32",
        config={"workload": "plain run_to_image", "rays": n, "detector": [128, 128], "seed": 0,
                "kernel_ms": kern_ms, "parity_x_bit_exact": parity},
        roofline={"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                  "frac": achieved / HBM_PEAK, "traffic": None},
        cpu_baseline={"value": n / cpu_s, "unit": "rays/s", "cores": 1, "kind": "port",
                      "sample": f"numpy oracle, full {n}-ray batch x {reps} reps"},
    )


WORKLOADS = {"plain": bench_plain}
DEFAULT = "plain"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default=DEFAULT, choices=sorted(WORKLOADS))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by scan position / detector tile with no exchange; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    res = WORKLOADS[args.workload](args.steps, args.warmup)
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
