"""Segmentation throughput on one MI355X.

One step = seeded initialisation + k-means to tolerance + connected
components + label compression over one synthetic image already resident in
HBM (split_hip_uniform_random_initialize + split_hip_segment).  Default
workload: shaded_4096 (4096x4096 Voronoi x Lambertian shading, k=32, stop at
centroid shift < 1e-4 or 100 iterations, 8-connectivity).

Prints one JSON line.  `roofline` is for the dominant kernel (the bounded
Lloyd iteration, one tiles_step launch that also finalises the previous
step): 12 bytes/pixel algorithmic traffic (three float planes) per iteration,
over its average time measured with HIP events on the stream it runs on, as
the slope of cluster() time between two fixed iteration counts (8 and 40).
`traffic` is the HBM bytes per bounded tiles_step launch of the same
iterations, from two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) of
`bench.py --profile-driver`, scaled by the counters of a 256 MiB copy.
`roofline.valu_issue` puts the same launches against the vector ALU: wave64
VALU instructions issued per launch (a third counter pass, SQ_INSTS_VALU)
over the launch's issue capacity (one per CU per cycle at 2.4 GHz).
`cpu_baseline` times the oracle's restatement of the reference pipeline on a
bounded crop of the same image on this host: on all OMP_NUM_THREADS threads
with the reference's host flags (`cores`), and on one thread of the checker
build (`single_thread`).

`--workload dropin_4096` times what a caller of the reference's C++ API gets
on the same image: uniform_random_initialize, cluster() without an image
width (the bounded path over rows of 16 points), connected_components and
compress_labels as separate calls, each host-synchronous.

`--workload batch_256` instead runs the batched path (256 images of uneven
size and k, one graph-replayed launch sequence) against the same images
through 256 separate segment calls.  Other workloads: cc_stress_8192 (CC +
compress of the spiral/checkerboard labels), segmeans_4096 (segment means +
propagation), color_4096 (Lab + beta features), ingest_4096 (host image
upload / download / 8-bit egress), erode_4096 (erosion of the segment
labels), and the image workloads of
split_amd.workloads (plain_512, stress_8192, ...).
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
CLOCK_HZ = 2.4e9   # MI355X peak engine clock
N_CU = 256


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="shaded_4096")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--cpu-crop", type=int, default=0,
                    help="side of the CPU baseline's crop (default: 2048 for k <= 64, "
                         "1024 above: about 10-30 s of CPU work)")
    ap.add_argument("--no-warm-start", action="store_true",
                    help="skip the learned warm-start comparison")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes behind roofline.traffic")
    ap.add_argument("--profile-driver", action="store_true",
                    help="run the fixed workload the counter passes profile and exit: "
                         "a 256 MiB device copy (traffic scale reference), the roofline "
                         "leg's 40-iteration tiled cluster(), CC and compress")
    return ap.parse_args()


ROOFLINE_LO, ROOFLINE_HI = 8, 40   # fixed iteration counts of the slope timing
COPY_BYTES = 256 << 20             # traffic scale reference


def profile_driver(w):
    """The fixed run the counter passes profile: a 256 MiB device copy (the
    scale reference), then exactly the roofline leg's longer cluster() run
    (seeded init, ROOFLINE_HI iterations, tiled path), then CC + compress."""
    import torch
    import split_amd as sa
    from split_amd import workloads as W
    ref = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    ref.fill_(1.0)
    ref.clone()
    torch.cuda.synchronize()
    width, height = w.image.width, w.image.height
    P = torch.from_numpy(W.make_image(w.image)).cuda()
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    L = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    sa.uniform_random_initialize(P, C, w.init_seed)
    it, _ = sa.cluster(P, C, L, ROOFLINE_HI, -1.0, width=width, height=height)
    sa.connected_components(L.view(height, width), S, connectivity=w.connectivity)
    n = sa.compress_labels(S)
    torch.cuda.synchronize()
    log(f"profile driver: {it} iterations, {n} segments")


def _counter_rows(exe, counter, workload, outdir):
    import csv
    import subprocess
    name = counter.lower()
    cmd = [exe, "--pmc", counter, "--output-format", "csv", "-d", str(outdir), "-o", name,
           "--", sys.executable, str(ROOT / "bench.py"), "--workload", workload,
           "--profile-driver"]
    env = dict(os.environ, TMPDIR="/tmp")
    subprocess.run(cmd, capture_output=True, text=True, timeout=180, env=env, check=True)
    files = sorted(Path(outdir).rglob(f"{name}_counter_collection.csv"))
    if not files:
        raise RuntimeError(f"no {name} counter file")
    with open(files[0], newline="") as f:
        recs = list(csv.DictReader(f))
    if recs and "Dispatch_Id" in recs[0]:  # in dispatch order
        recs.sort(key=lambda r: int(r["Dispatch_Id"]))
    return [(r["Kernel_Name"], float(r["Counter_Value"])) for r in recs]


def measure_traffic(workload):
    """HBM bytes per bounded Lloyd launch over the roofline leg's iterations,
    from two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) of the profile
    driver, each scaled by what the counter reads for the 256 MiB copy."""
    import shutil
    import tempfile
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    res = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        for counter in ("FETCH_SIZE", "WRITE_SIZE", "SQ_INSTS_VALU"):
            rows = _counter_rows(exe, counter, workload, Path(d) / counter.lower())
            copy = [v for k, v in rows if "copyBuffer" in k]
            # the bounded launches: tiles_step<CPL, MODE, false>; for k > 64
            # each step is enqueued twice (box-pruned, then ball walk; one of
            # them returns after the finalisation) and the pair is one step
            box = [(k.startswith("void splithip::tiles_step<") and "false>" in k, v)
                   for k, v in rows]
            ball = [k.startswith("void splithip::tiles_step_ball<") for k, _ in rows]
            steps = []
            for i, (is_box, v) in enumerate(box):
                if not is_box:
                    continue
                if i + 1 < len(rows) and ball[i + 1]:
                    v += rows[i + 1][1]
                steps.append(v)
            # launches 2 .. HI run bounded steps; launch HI + 1 only finalises
            want = ROOFLINE_HI
            if not copy or len(steps) != want:
                raise RuntimeError(f"{counter}: {len(copy)} copies, {len(steps)} bounded launches "
                                   f"(want {want})")
            sel = steps[ROOFLINE_LO - 1:ROOFLINE_HI - 1]   # iterations LO+1 .. HI
            if counter == "SQ_INSTS_VALU":
                res[counter] = sum(sel) / len(sel)
            else:
                scale = COPY_BYTES / (max(copy) * 1024.0)
                res[counter] = (sum(sel) / len(sel) * 1024.0 * scale, scale)
    (rd, rs), (wr, ws) = res["FETCH_SIZE"], res["WRITE_SIZE"]
    return {"read": rd, "write": wr, "read_scale": rs, "write_scale": ws,
            "valu_insts": res["SQ_INSTS_VALU"]}


def batch_bench(args):
    """Batched segmentation of batch_256 (one graph-replayed launch sequence)
    versus the same images through 256 separate split_hip_segment calls."""
    import torch
    import split_amd as sa
    from split_amd import workloads as W
    ws = W.batch_workloads()
    total = sum(w.n for w in ws)
    log(f"bench: generating {len(ws)} images, {total / 1e6:.1f} Mpix")
    torch.cuda.init()
    ktot = sum(w.k for w in ws)
    cflat = torch.empty(3 * ktot, dtype=torch.float32, device="cuda")
    images, off = [], 0
    for i, w in enumerate(ws):
        P = torch.from_numpy(W.make_image(w.image)).cuda()
        C = cflat[off:off + 3 * w.k].view(3, w.k)
        off += 3 * w.k
        sa.uniform_random_initialize(P, C, w.init_seed)
        images.append(dict(points=P, width=w.image.width, height=w.image.height, centroids=C,
                           labels=torch.empty(w.n, dtype=torch.int32, device="cuda"),
                           segments=torch.empty(w.n, dtype=torch.int32, device="cuda")))
        if i % 32 == 31:
            log(f"bench: {i + 1} images resident")
    init = cflat.clone()
    temp = torch.empty(sa.segment_batch_workspace(images), dtype=torch.uint8, device="cuda")
    stemp = torch.empty(max(sa.segment_workspace(w.image.width, w.image.height, w.k)
                            for w in ws), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    w0 = ws[0]

    def batched():
        cflat.copy_(init)
        return sa.segment_batch(images, max_iter=w0.max_iter, threshold=w0.threshold,
                                connectivity=w0.connectivity, temp=temp)

    def separate():
        cflat.copy_(init)
        return [sa.segment(im["points"], im["width"], im["height"], im["centroids"],
                           im["labels"], im["segments"], max_iter=w0.max_iter,
                           threshold=w0.threshold, connectivity=w0.connectivity, temp=stemp)
                for im in images]

    def timed(fn, steps):
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            r = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps, r

    for i in range(args.warmup):
        rb = batched()
        log(f"bench: batched warmup {i}: iters min/max "
            f"{min(r.iterations for r in rb)}/{max(r.iterations for r in rb)}")
    ms_b, rb = timed(batched, args.steps)
    log(f"bench: batched {ms_b:.2f} ms/step")
    separate()
    ms_s, rs = timed(separate, max(1, args.steps))
    log(f"bench: separate {ms_s:.2f} ms/step")
    same = all(a.iterations == b.iterations and a.num_segments == b.num_segments
               for a, b in zip(rb, rs))
    value = total / (ms_b / 1e3) / 1e6

    # roofline of the batched bounded Lloyd step (batch_step_bounded<false>):
    # every image runs a fixed count of iterations (threshold < 0, no graph),
    # and the slope between 8 and 24 iterations is the time of one launch
    # over all 256 images; algorithmic bytes: 12 B per pixel of the batch
    def fixed(iters_):
        cflat.copy_(init)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        sa.segment_batch(images, max_iter=iters_, threshold=-1.0,
                         connectivity=w0.connectivity, use_graph=False, temp=temp)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    fixed(8)
    lo, hi = 8, 24
    t_lo = min(fixed(lo) for _ in range(2))
    t_hi = min(fixed(hi) for _ in range(2))
    per_launch_ms = (t_hi - t_lo) / (hi - lo)
    achieved = 12 * total / (per_launch_ms / 1e3) / 1e9
    roof = {
        "kernel": "batched Lloyd iteration: batch_step_bounded<false> over all 256 images "
                  "(finalise each image, skip test, re-assign, exact sums)",
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "per_launch_us": round(per_launch_ms * 1e3, 2), "bytes_per_launch": 12 * total,
        "timing": f"slope of segment_batch time between {lo} and {hi} fixed iterations "
                  f"(HIP events, no graph)",
    }
    cpu = None
    if not args.no_cpu_baseline:
        import numpy as np
        import oracle
        log("bench: cpu baseline (first images of the batch, until ~15 s)")
        ref = oracle.lib_reference_flags()
        done_px, spent, used = 0, 0.0, 0
        for i, w in enumerate(ws):
            if spent > 15.0:
                break
            img = images[i]["points"].cpu().numpy()
            init_i = init[sum(x.k for x in ws[:i]) * 3:sum(x.k for x in ws[:i + 1]) * 3]
            init_i = init_i.view(3, w.k).cpu().numpy()
            dt, _ = _cpu_seconds(lambda: cpu_pipeline(img, w.image.width, w.image.height,
                                                      init_i, rb[i].iterations, w.connectivity,
                                                      handle=ref))
            done_px += w.n
            spent += dt
            used += 1
        cpu = {"value": round(done_px / spent / 1e6, 4), "unit": "Mpix/s",
               "cores": oracle.threads(handle=ref), "kind": "port",
               "sample": (f"the first {used} images of the batch ({done_px / 1e6:.1f} Mpix), "
                          f"each with its GPU iteration count: the reference pipeline restated "
                          f"(dense float distances, label-sorted reduce_by_key centroids, "
                          f"union-find CC, sort/unique compress), oracle C with the reference's "
                          f"flags, assignment threaded: {spent:.2f} s")}
    out = {
        "metric": "segmented Mpix/s (k-means+CC) & Lloyd iters to tol, warm-start vs plain",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_b, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {
            "workload": "batch_256",
            "images": len(ws),
            "pixels": total,
            "k": "per image from {4,8,16,32,64}",
            "init": "seeded sample (plain, no warm start)",
            "stop": f"shift<{w0.threshold} or {w0.max_iter} iters",
            "connectivity": w0.connectivity,
            "lloyd_iters_mean": round(sum(r.iterations for r in rb) / len(rb), 2),
            # sum over the images of pixels x iterations (the bounded launches
            # of a run process sum(n_i (I_i - 1)) pixel-iterations)
            "pixel_iterations": sum(w.n * r.iterations for w, r in zip(ws, rb)),
            "parallelism": "single GPU",
        },
        "separate_launches": {"ms_per_step": round(ms_s, 4),
                              "value": round(total / (ms_s / 1e3) / 1e6, 3),
                              "speedup_batched": round(ms_s / ms_b, 3),
                              "same_iterations_and_segments": same},
        "roofline": roof,
    }
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


def cc_bench(args):
    """Connected components (8-connectivity) + compress_labels of the
    8192x8192 spiral/checkerboard stress label image."""
    import torch
    import split_amd as sa
    from split_amd import workloads as W
    side = 8192
    log(f"bench: generating the {side}x{side} spiral/checkerboard labels")
    lab = torch.from_numpy(W.make_cc_stress(side, side)).cuda()
    out = torch.empty(side * side, dtype=torch.int32, device="cuda")
    cc_temp = torch.empty(max(1, sa.lib().split_hip_connected_components_workspace(side, side)),
                          dtype=torch.uint8, device="cuda")
    cp_temp = torch.empty(sa.lib().split_hip_compress_labels_workspace(side * side,
                                                                       side * side),
                          dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        sa.connected_components(lab, out, connectivity=8, temp=cc_temp)
        return sa.compress_labels(out, temp=cp_temp)

    for _ in range(args.warmup):
        nseg = step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        nseg = step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    n = side * side
    # the two calls apart: connected_components (cc_clear, cc_tile, cc_merge,
    # cc_jump, cc_final: labels in, components out = 8 B/px algorithmic) and
    # compress_labels (5 kernels, values in and ranks out = 8 B/px)
    ms_cc = _events_ms(lambda: sa.connected_components(lab, out, connectivity=8,
                                                       temp=cc_temp), args.steps, 1)
    ms_cp = ms - ms_cc
    roof = {
        "kernel": "connected_components call: cc_clear, cc_tile<8,64>, cc_merge<8,64>, "
                  "cc_jump, cc_final<64> (cc_jump dominates on this image: root chains "
                  "along the spiral arms)",
        "bound": "hbm",
        "achieved": round(8 * n / (ms_cc / 1e3) / 1e9, 1),
        "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(8 * n / (ms_cc / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "traffic": None,
        "per_call_us": round(ms_cc * 1e3, 1),
        "bytes_per_call": 8 * n,
        "compress_labels_call": {"per_call_us": round(ms_cp * 1e3, 1),
                                 "achieved": round(8 * n / (ms_cp / 1e3) / 1e9, 1),
                                 "frac": round(8 * n / (ms_cp / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
    }
    cpu = None
    if not args.no_cpu_baseline:
        import oracle
        log("bench: cpu baseline (union-find CC + compress of the same labels)")
        ref = oracle.lib_reference_flags()
        host = lab.cpu().numpy()
        dt, (rep_, _) = _cpu_seconds(lambda: oracle.cc_uf(host, 8, handle=ref))
        dtc, _ = _cpu_seconds(lambda: oracle.compress(rep_, handle=ref))
        cpu = {"value": round(n / (dt + dtc) / 1e6, 3), "unit": "Mpix/s", "cores": 1,
               "kind": "port",
               "sample": (f"the whole {side}x{side} label image: union-find CC (8-conn, the "
                          f"reference's converged result) + sort/unique compress, oracle C "
                          f"with the reference's flags, one thread (serial code): "
                          f"{dt + dtc:.2f} s")}
    out_line = {
        "metric": "connected components + compress Mpix/s (8-conn, stress labels)",
        "value": round(n / (ms / 1e3) / 1e6, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "cc_stress_8192", "image": f"{side}x{side}",
                   "connectivity": 8, "segments": nseg, "parallelism": "single GPU"},
        "algorithmic_bytes_per_step": 16 * n,
        "hbm_frac": round(16 * n / (ms / 1e3) / (HBM_PEAK_GBS * 1e9), 4),
        "roofline": roof,
    }
    if cpu:
        out_line["cpu_baseline"] = cpu
    print(json.dumps(out_line), flush=True)


def segmeans_bench(args):
    """Segment means of the shaded_4096 components: calculate_centroids with
    k = number of segments, then propagate_centroids (the consumer of CC +
    compress in the reference's sample)."""
    import torch
    import split_amd as sa
    from split_amd import workloads as W
    w = W.SHADED_4096
    P = torch.from_numpy(W.make_image(w.image)).cuda()
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    L = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    sa.uniform_random_initialize(P, C, w.init_seed)
    r = sa.segment(P, 4096, 4096, C, L, S, max_iter=w.max_iter, threshold=w.threshold,
                   connectivity=w.connectivity)
    k = r.num_segments
    M = torch.empty((3, k), dtype=torch.float32, device="cuda")
    out = torch.empty_like(P)
    temp = torch.empty(sa.lib().split_hip_calculate_centroids_workspace(w.n, k),
                       dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        sa.calculate_centroids(S, P, M, temp=temp)
        sa.propagate_centroids(S, M, out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print(json.dumps({
        "metric": "segment means + propagation Mpix/s",
        "value": round(w.n / (ms / 1e3) / 1e6, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "segmeans_4096", "image": "4096x4096", "segments": k,
                   "parallelism": "single GPU"},
        "algorithmic_bytes_per_step": 32 * w.n,
        "hbm_frac": round(32 * w.n / (ms / 1e3) / (HBM_PEAK_GBS * 1e9), 4),
    }), flush=True)


def _events_ms(fn, steps, warmup):
    """Average ms of fn() over `steps` calls on the current stream (HIP
    events), after `warmup` calls."""
    import torch
    stream = torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def color_bench(args):
    """The features k-means clusters in the reference's sample (main.cu:216-225):
    rgb -> Lab, copy of L, beta over L, on the shaded_4096 image.  The fused
    split_hip_lab_beta_feature against the separate calls (convert, copy,
    beta_feature)."""
    import torch
    import split_amd as sa
    from split_amd import workloads as W
    w = W.SHADED_4096
    n = w.n
    P = torch.from_numpy(W.make_image(w.image)).cuda()
    F = torch.empty_like(P)
    lum = torch.empty(n, dtype=torch.float32, device="cuda")
    temp = torch.empty(sa.lib().split_hip_beta_feature_workspace(), dtype=torch.uint8,
                       device="cuda")
    ms_fused = _events_ms(lambda: sa.lab_beta_feature(P, F, lum, 1e5, temp=temp),
                          args.steps * 10, args.warmup)

    def separate():
        sa.convert_color_space(P, F, "rgb_to_lab")
        lum.copy_(F[0])
        sa.beta_feature(F, F[0], 1e5, temp=temp)

    ms_sep = _events_ms(separate, args.steps * 10, args.warmup)
    algo = 28 * n  # read 3 planes, write 3 feature planes and the luminance copy
    print(json.dumps({
        "metric": "Lab + beta feature Mpix/s (the k-means input features)",
        "value": round(n / (ms_fused / 1e3) / 1e6, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps * 10, "warmup": args.warmup, "ms_per_step": round(ms_fused, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "color_4096", "image": "4096x4096", "mu": 1e5,
                   "parallelism": "single GPU"},
        "algorithmic_bytes_per_step": algo,
        "hbm_frac": round(algo / (ms_fused / 1e3) / (HBM_PEAK_GBS * 1e9), 4),
        "designed_bytes_per_step": 44 * n,
        "separate_calls": {"ms_per_step": round(ms_sep, 4),
                           "value": round(n / (ms_sep / 1e3) / 1e6, 3),
                           "speedup_fused": round(ms_sep / ms_fused, 3)},
    }), flush=True)


def erode_bench(args):
    """Erosion (n = 10, the sample's call at samples/src/main.cu:302) of the
    shaded_4096 segment labels: the column sweep then the row sweep."""
    import torch
    import split_amd as sa
    from split_amd import workloads as W
    w = W.SHADED_4096
    P = torch.from_numpy(W.make_image(w.image)).cuda()
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    L = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    sa.uniform_random_initialize(P, C, w.init_seed)
    r = sa.segment(P, 4096, 4096, C, L, S, max_iter=w.max_iter, threshold=w.threshold,
                   connectivity=w.connectivity)
    E = torch.empty_like(S).view(4096, 4096)
    temp = torch.empty(sa.lib().split_hip_erode_workspace(4096, 4096), dtype=torch.uint8,
                       device="cuda")
    stream = torch.cuda.current_stream()
    total, reps = 0.0, args.steps * 10
    for i in range(args.warmup + reps):
        E.view(-1).copy_(S)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        sa.erode(E, 10, temp=temp)
        e1.record(stream)
        torch.cuda.synchronize()
        if i >= args.warmup:
            total += e0.elapsed_time(e1)
    ms = total / reps
    kept = int((E >= 0).sum().item())
    print(json.dumps({
        "metric": "erosion Mpix/s (n = 10, segment labels)",
        "value": round(w.n / (ms / 1e3) / 1e6, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": reps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "erode_4096", "image": "4096x4096", "niterations": 10,
                   "segments": r.num_segments, "pixels_kept": kept,
                   "parallelism": "single GPU"},
        "algorithmic_bytes_per_step": 8 * w.n,
        "hbm_frac": round(8 * w.n / (ms / 1e3) / (HBM_PEAK_GBS * 1e9), 4),
    }), flush=True)


def ingest_bench(args):
    """Host ingest / egress of the shaded_4096 image: interleaved pinned host
    floats -> device planes (one PCIe copy + device deinterleave), planes ->
    host floats, and the 8-bit egress (device quantise + 3 B/px copy back).
    The reference's call pattern (one 2-D copy of 4-byte rows per channel,
    samples/src/main.cu:84-114) is timed on a 256x256 crop."""
    import ctypes
    import torch
    import split_amd as sa
    from split_amd import workloads as W
    w = W.SHADED_4096
    n = w.n
    img = W.make_image(w.image)
    host = torch.from_numpy(img.T.copy()).pin_memory()          # (N, 3) interleaved
    back = torch.empty_like(host).pin_memory()
    q_host = torch.empty((n, 3), dtype=torch.uint8).pin_memory()
    P = torch.empty((3, n), dtype=torch.float32, device="cuda")
    Q = torch.empty((n, 3), dtype=torch.uint8, device="cuda")
    temp = torch.empty(sa.lib().split_hip_interleaved_workspace(n, 3), dtype=torch.uint8,
                       device="cuda")
    reps = max(3, args.steps)
    ms_up = _events_ms(lambda: sa.make_device_image(host, P, temp=temp), reps, args.warmup)
    ok_up = bool(torch.equal(P.cpu(), host.t().contiguous()))
    ms_down = _events_ms(lambda: sa.make_host_image(P, back, temp=temp), reps, args.warmup)
    ok_down = bool(torch.equal(back, host))

    def egress8():
        sa.quantize(P, Q)
        q_host.copy_(Q, non_blocking=True)
        torch.cuda.synchronize()

    ms_q = _events_ms(egress8, reps, args.warmup)
    ms_dd = _events_ms(lambda: sa.deinterleave(temp.view(torch.float32)[:3 * n].view(n, 3), P),
                       reps * 10, args.warmup)
    # the reference's pattern on a 256x256 crop: per channel, one 2-D copy of
    # npixels rows of one float with a source pitch of 3 floats
    side = 256
    m = side * side
    small = host[:m].contiguous().pin_memory()
    Ps = torch.empty((3, m), dtype=torch.float32, device="cuda")
    # the HIP runtime torch already loaded (same file: no second runtime)
    with open("/proc/self/maps") as f:
        hip_path = next(ln.split()[-1] for ln in f if "libamdhip64.so" in ln)
    hip = ctypes.CDLL(hip_path)
    hip.hipMemcpy2D.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                ctypes.c_int]

    def ref_pattern():
        for c in range(3):
            rc = hip.hipMemcpy2D(Ps.data_ptr() + 4 * m * c, 4, small.data_ptr() + 4 * c, 12,
                                 4, m, 1)
            if rc != 0:
                raise RuntimeError(f"hipMemcpy2D failed: {rc}")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref_pattern()
    torch.cuda.synchronize()
    ms_ref = (time.perf_counter() - t0) * 1e3
    ok_ref = bool(torch.equal(Ps.cpu(), small.t().contiguous()))
    gbs = lambda b, ms: round(b / (ms / 1e3) / 1e9, 2)
    print(json.dumps({
        "metric": "image ingest Mpix/s (interleaved pinned host floats -> device planes)",
        "value": round(n / (ms_up / 1e3) / 1e6, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": reps, "warmup": args.warmup, "ms_per_step": round(ms_up, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator)",
        "config": {"workload": "ingest_4096", "image": "4096x4096", "channels": 3,
                   "parallelism": "single GPU"},
        "upload": {"ms": round(ms_up, 4), "pcie_GB_s": gbs(12 * n, ms_up), "exact": ok_up},
        "download": {"ms": round(ms_down, 4), "pcie_GB_s": gbs(12 * n, ms_down),
                     "exact": ok_down},
        "egress_u8": {"ms": round(ms_q, 4), "Mpix_s": round(n / (ms_q / 1e3) / 1e6, 3)},
        "device_deinterleave": {"ms": round(ms_dd, 4), "hbm_GB_s": gbs(24 * n, ms_dd),
                                "hbm_frac": round(24 * n / (ms_dd / 1e3) / (HBM_PEAK_GBS * 1e9),
                                                  4)},
        "reference_pattern_256": {"ms": round(ms_ref, 3),
                                  "Mpix_s": round(m / (ms_ref / 1e3) / 1e6, 3),
                                  "ours_same_size_Mpix_s_est": round(n / (ms_up / 1e3) / 1e6, 3),
                                  "exact": ok_ref},
    }), flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _cpu_seconds(fn):
    t0 = time.perf_counter()
    r = fn()
    return time.perf_counter() - t0, r


def cpu_pipeline(planes, width, height, init, iters, conn, handle=None):
    """The reference pipeline restated on the CPU (oracle): iters Lloyd
    iterations of dense float distances + label-sorted reduce_by_key
    centroids, union-find CC, sort/unique compress; returns the labels."""
    import oracle
    _, labels, it, _ = oracle.ref_cluster(planes, init, iters, -1.0, handle=handle)
    rep, _ = oracle.cc_uf(labels.reshape(height, width), conn, handle=handle)
    oracle.compress(rep, handle=handle)
    return labels


def cpu_baseline(img, width, height, w, gpu_iters, crop):
    """Reference pipeline restated on the CPU (oracle) on a center crop, on
    this host: all threads with the reference's host flags
    (-Ofast -march=native -fopenmp, split's common.pri:27-32; OpenMP splits
    the per-pixel assignment, as the reference's OMP Thrust backend would;
    the counting sort, CC and compress run on one thread), and one thread
    of the checker build (-O3, no fast-math)."""
    import numpy as np
    import oracle
    c = min(crop, width, height)
    y0, x0 = (height - c) // 2, (width - c) // 2
    planes = img.reshape(3, height, width)[:, y0:y0 + c, x0:x0 + c]
    sub = np.ascontiguousarray(planes.reshape(3, c * c))
    idx = oracle.sample_indices(w.init_seed, c * c, w.k)
    init = np.ascontiguousarray(sub[:, idx])
    ref = oracle.lib_reference_flags()
    dt1, lab1 = _cpu_seconds(lambda: cpu_pipeline(sub, c, c, init, gpu_iters, w.connectivity))
    dtp, labp = _cpu_seconds(lambda: cpu_pipeline(sub, c, c, init, gpu_iters, w.connectivity,
                                                  handle=ref))
    cores = oracle.threads(handle=ref)
    agree = float(np.mean(lab1 == labp))
    return {
        "value": round(c * c / dtp / 1e6, 4),
        "unit": "Mpix/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"center {c}x{c} crop of the {w.name} image, k={w.k}, {gpu_iters} Lloyd "
                   f"iterations of the reference algorithm (dense float distances, "
                   f"label-sorted reduce_by_key centroids) + union-find CC "
                   f"({w.connectivity}-conn) + sort/unique compress; oracle C built with the "
                   f"reference's flags {' '.join(oracle.REFERENCE_FLAGS)}, {cores} threads "
                   f"(assignment threaded, sort/CC/compress on one thread): {dtp:.2f} s; "
                   f"checker build (-O3, no fast-math), 1 thread: {dt1:.2f} s; "
                   f"labels agree on {agree * 100:.3f}% of pixels"),
        "single_thread": {"value": round(c * c / dt1 / 1e6, 4), "cores": 1,
                          "seconds": round(dt1, 2)},
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the segmentation path runs on one GPU here; "
                                   "--gpus must be 1 (multi-GPU sharding is not "
                                   "implemented)"}))
        sys.exit(2)

    import numpy as np
    import torch

    import split_amd as sa
    from split_amd import workloads as W

    if args.workload == "batch_256":
        batch_bench(args)
        return
    if args.workload == "cc_stress_8192":
        cc_bench(args)
        return
    if args.workload == "segmeans_4096":
        segmeans_bench(args)
        return
    if args.workload == "color_4096":
        color_bench(args)
        return
    if args.workload == "ingest_4096":
        ingest_bench(args)
        return
    if args.workload == "erode_4096":
        erode_bench(args)
        return
    dropin = args.workload == "dropin_4096"
    w = W.SHADED_4096 if dropin else W.WORKLOADS[args.workload]
    if args.profile_driver:
        profile_driver(w)
        return
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

    if dropin:
        # the reference's call sequence (samples/src/main.cu:228-253) through
        # the C ABI the C++ drop-ins forward to, with their temp contract
        dtemp = torch.empty(max(sa.lib().split_hip_cluster_workspace(n, w.k, 3),
                                sa.lib().split_hip_connected_components_workspace(width,
                                                                                  height),
                                sa.lib().split_hip_compress_labels_workspace(n, n)),
                            dtype=torch.uint8, device="cuda")

        def step():
            sa.uniform_random_initialize(P, C, w.init_seed)
            it, shift = sa.cluster(P, C, L, w.max_iter, w.threshold, temp=dtemp)
            sa.connected_components(L.view(height, width), S, connectivity=w.connectivity,
                                    temp=dtemp)
            ns = sa.compress_labels(S, temp=dtemp)
            return sa.SegmentResult(it, shift, ns)

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
            "workload": args.workload,
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
            geo = {} if dropin else dict(width=width, height=height)
            sa.cluster(P, Ct, Lt, iters_, -1.0, temp=ctemp, **geo)
            b.record(stream)
            torch.cuda.synchronize()
            return a.elapsed_time(b)

        timed(4)
        lo, hi = ROOFLINE_LO, ROOFLINE_HI
        t_lo = min(timed(lo) for _ in range(3))
        t_hi = min(timed(hi) for _ in range(3))
        per_launch_ms = (t_hi - t_lo) / (hi - lo)
        bytes_per_launch = 12 * n
        achieved = bytes_per_launch / (per_launch_ms / 1e3) / 1e9
        kname = ("tiles_step<1, 3, false> (finalise the previous step, skip test, "
                 "re-assignment: whole tiles box-pruned, queued pixels against "
                 "every centroid, exact sums)" if w.k <= 64 else
                 "tiles_step_ball<3> (finalise the previous step, skip test, "
                 "neighbour-ball re-assignment, exact sums)")
        out["roofline"] = {
            "kernel": f"Lloyd iteration: {kname}",
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
        if dropin:
            out["roofline"]["traffic_note"] = ("counter passes run on the default "
                                               "workload's image-tiled launches")
        if not args.no_traffic and not dropin:
            log("bench: counter passes for roofline.traffic")
            try:
                t = measure_traffic(w.name)
            except Exception as e:  # the profiler is optional; say why it is absent
                t = None
                out["roofline"]["traffic_error"] = str(e)[:200]
            if t is not None:
                out["roofline"]["traffic"] = round(t["read"] + t["write"])
                out["roofline"]["traffic_read"] = round(t["read"])
                out["roofline"]["traffic_write"] = round(t["write"])
                out["roofline"]["traffic_scale"] = [round(t["read_scale"], 4),
                                                    round(t["write_scale"], 4)]
                # the vector ALU: wave64 VALU instructions issued per launch
                # (SQ_INSTS_VALU) against one per CU per cycle (4 SIMDs, a
                # wave64 instruction every 4 cycles each) over the launch
                cap = per_launch_ms / 1e3 * CLOCK_HZ * N_CU
                out["roofline"]["valu_issue"] = {
                    "insts_per_launch": round(t["valu_insts"]),
                    "capacity_per_launch": round(cap),
                    "frac": round(t["valu_insts"] / cap, 4),
                }
                log(f"bench: traffic {t['read'] / 1e6:.1f} MB read + {t['write'] / 1e6:.1f} MB "
                    f"written per launch")

    if not args.no_warm_start and args.workload == "shaded_4096":
        # The learned warm start against the plain method on its evaluation
        # set (256 config-2-style 1024x1024 images, k=16): plain separate
        # segment() calls, plain batched, the network's prediction, the
        # prediction refined on the 256^2 coarse image (the shipped warm
        # start) and, as the ablation, the seeded sample refined the same way.
        from split_amd import warmstart as WS
        log("bench: warm-start evaluation (256 x 1024^2, k=16)")
        ev = WS.evaluate(WS.WarmStartNet.load(), count=256, side=1024, separate=True,
                         log=log)["summary"]
        ws = {
            "workload": "256 x 1024x1024 shaded Voronoi, k=16, shift<1e-4 or 100 iters",
            "plain_separate_mpix_s": round(ev["separate_plain_mpix_s"], 1),
            "plain_batched_mpix_s": round(ev["plain_mpix_s"], 1),
            "plain_iters_mean": round(ev["plain_iters_mean"], 2),
            "plain_cap_hits": ev["plain_cap_hits"],
        }
        for name in ("warm", "warm_coarse", "plain_coarse"):
            ws.update({
                f"{name}_batched_mpix_s": round(ev[f"{name}_mpix_s"], 1),
                f"{name}_iters_mean": round(ev[f"{name}_iters_mean"], 2),
                f"{name}_cap_hits": ev[f"{name}_cap_hits"],
                f"{name}_inertia_ratio": round(ev[f"{name}_inertia_ratio_total"], 4),
            })
            if f"{name}_coarse_stage_iters_mean" in ev:
                ws[f"{name}_coarse_stage_iters_mean"] = round(
                    ev[f"{name}_coarse_stage_iters_mean"], 2)
                ws[f"{name}_pixel_iters_mean"] = round(ev[f"{name}_pixel_iters_mean"], 2)
        ws["speedup_warm_coarse_batched_vs_plain_separate"] = round(
            ev["speedup_batched_warm_coarse_vs_separate_plain"], 3)
        ws["speedup_warm_coarse_vs_plain_batched"] = round(
            ev["plain_ms"] / ev["warm_coarse_ms"], 3)
        out["warm_start"] = ws

    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        crop = args.cpu_crop or (2048 if w.k <= 64 else 1024)
        out["cpu_baseline"] = cpu_baseline(img, width, height, w, iters[-1], crop)

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
