"""Segmentation throughput on one MI355X.

One step = seeded initialisation + k-means to tolerance + connected
components + label compression over one synthetic image already resident in
HBM (split_hip_uniform_random_initialize + split_hip_segment).  Default
workload: shaded_4096 (4096x4096 Voronoi x Lambertian shading, k=32, stop at
centroid shift < 1e-4 or 100 iterations, 8-connectivity).

Prints one JSON line.  `roofline` is for the dominant kernel (the fused Lloyd
step): 12 bytes/pixel algorithmic traffic (three float planes) per launch,
over its average launch time measured with HIP events on the stream it runs
on, as the slope of cluster() time between two fixed iteration counts.
`cpu_baseline` times the oracle's restatement of the reference pipeline on a
bounded crop of the same image, single-threaded, on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="shaded_4096")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--cpu-crop", type=int, default=1024)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(img, width, height, w, gpu_iters, crop):
    """Reference pipeline restated on the CPU (oracle), on a crop."""
    import numpy as np
    import oracle
    c = min(crop, width, height)
    y0, x0 = (height - c) // 2, (width - c) // 2
    planes = img.reshape(3, height, width)[:, y0:y0 + c, x0:x0 + c]
    sub = np.ascontiguousarray(planes.reshape(3, c * c))
    idx = oracle.sample_indices(w.init_seed, c * c, w.k)
    init = np.ascontiguousarray(sub[:, idx])
    t0 = time.perf_counter()
    cent, labels, it, _ = oracle.ref_cluster(sub, init, gpu_iters, -1.0)
    rep, _ = oracle.cc_uf(labels.reshape(c, c), w.connectivity)
    oracle.compress(rep)
    dt = time.perf_counter() - t0
    return {
        "value": round(c * c / dt / 1e6, 4),
        "unit": "Mpix/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"center {c}x{c} crop of the {w.name} image, k={w.k}, {it} Lloyd "
                   f"iterations of the reference algorithm (dense float distances, "
                   f"sorted reduce_by_key centroids) + union-find CC "
                   f"({w.connectivity}-conn) + sort/unique compress; oracle C, 1 thread; "
                   f"{dt:.2f} s"),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(see DESIGN.md for how the path shards)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import split_amd as sa
    from split_amd import workloads as W

    w = W.WORKLOADS[args.workload]
    width, height, n = w.image.width, w.image.height, w.n
    log(f"bench: generating {w.name} ({width}x{height})")
    img = W.make_image(w.image)
    torch.cuda.init()
    P = torch.from_numpy(img).cuda()
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    L = torch.empty(n, dtype=torch.int32, device="cuda")
    S = torch.empty(n, dtype=torch.int32, device="cuda")
    temp = torch.empty(sa.segment_workspace(width, height, w.k), dtype=torch.uint8,
                       device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        sa.uniform_random_initialize(P, C, w.init_seed)
        return sa.segment(P, width, height, C, L, S, max_iter=w.max_iter,
                          threshold=w.threshold, connectivity=w.connectivity, temp=temp)

    for i in range(args.warmup):
        r = step()
        log(f"bench: warmup {i}: iters={r.iterations} segments={r.num_segments}")
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    iters = []
    for _ in range(args.steps):
        r = step()
        iters.append(r.iterations)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    log(f"bench: {args.steps} steps, {ms:.3f} ms/step (events), wall {wall / args.steps * 1e3:.3f}"
        f" ms/step, iters={iters}, segments={r.num_segments}")
    value = n / (ms / 1e3) / 1e6

    out = {
        "metric": "segmented Mpix/s (k-means+CC) & Lloyd iters to tol, warm-start vs plain",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": w.name,
            "image": f"{width}x{height}",
            "k": w.k,
            "init": "seeded sample (plain, no warm start)",
            "stop": f"shift<{w.threshold} or {w.max_iter} iters",
            "connectivity": w.connectivity,
            "lloyd_iters": iters[-1],
            "segments": r.num_segments,
            "parallelism": "single GPU",
        },
    }

    if not args.no_roofline:
        # Dominant kernel: the fused Lloyd step.  Average launch time as the
        # slope of cluster() time between two fixed iteration counts (same
        # stream, HIP events), which removes the fixed set-up/label-pass cost.
        Ct = torch.empty_like(C)
        Lt = torch.empty_like(L)
        ctemp = torch.empty(sa.lib().split_hip_cluster_workspace(n, w.k, 3),
                            dtype=torch.uint8, device="cuda")

        def timed(iters_):
            sa.uniform_random_initialize(P, Ct, w.init_seed)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            sa.cluster(P, Ct, Lt, iters_, -1.0, temp=ctemp, width=width, height=height)
            b.record(stream)
            torch.cuda.synchronize()
            return a.elapsed_time(b)

        timed(4)
        lo, hi = 8, 40
        t_lo = min(timed(lo) for _ in range(3))
        t_hi = min(timed(hi) for _ in range(3))
        per_launch_ms = (t_hi - t_lo) / (hi - lo)
        bytes_per_launch = 12 * n
        achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9
        out["roofline"] = {
            "kernel": "lloyd_step (assign + accumulate + finalize)",
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "per_launch_us": round(per_launch_ms * 1e3, 2),
            "bytes_per_launch": bytes_per_launch,
        }
        log(f"bench: lloyd_step {per_launch_ms * 1e3:.2f} us/launch -> {achieved:.1f} GB/s")

    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        out["cpu_baseline"] = cpu_baseline(img, width, height, w, iters[-1], args.cpu_crop)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
