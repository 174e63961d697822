"""Learned warm start: a small CNN that predicts 16 starting centroids.

Host side of the gfx950 network kernels (split_amd/csrc/warmstart.hip,
C ABI in include/split_hip.h, split_hip_ws_*).  The reference's cluster()
takes its initial centroids as an in/out argument
(split/include/split/device/kmeans/cluster.cuh:19-22,31-37) and its caller
fills them by sampling pixels (samples/src/main.cu:230-233); this module
fills them with the network's prediction instead.

    net = WarmStartNet.load()                       # committed weights
    imgs = [dict(points=P, width=w, height=h, centroids=C)]       # C (3, 16)
    net.predict_refined(imgs)   # prediction, refined on the 256^2 resample
    net.predict(imgs)           # the prediction alone

The refinement runs the batched clustering (connectivity 0) on the same
area-average resample the network reads, as float planes, from the
prediction until shift < 1e-4: a coarse-to-fine start whose iterations cost
1/16 of a full-size one on a 1024^2 image.

Training (on the GPU, by the project's own forward/backward kernels):

    python -m split_amd.warmstart train --images 50000 --epochs 12

generates config-2-style 256x256 images on the device, clusters each with
plain k-means (seeded-sample init, k=16, shift < 1e-4 or 100 iterations)
through the batched path, sorts the converged centroids by luminance as
targets, and fits the network with Adam on the MSE of its sigmoid outputs.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import (ImageDesc, _check, _dev, _points, _stream, _torch, lib, segment_batch,
               segment_batch_workspace, uniform_random_initialize)
from . import workloads as W

K = 16
SIDE = 256
OUT = 3 * K
WEIGHTS = Path(__file__).resolve().parent / "weights" / "warmstart_k16.npz"
LOSS_KINDS = {"mse": 0, "chamfer": 1}

# conv1..4, dense5, dense6: (Cin, Cout, K); Kp = roundup(K + 1, 32)
_LAYERS = [(3, 16, 27), (16, 32, 144), (32, 64, 288), (64, 64, 576), (64, 128, 64),
           (128, OUT, 128)]


def _kp(k: int) -> int:
    return (k + 1 + 31) // 32 * 32


def param_count() -> int:
    return sum(cout * _kp(kk) for _, cout, kk in _LAYERS)


def init_params(seed: int) -> np.ndarray:
    """He-normal weights, zero biases, in the kernels' parameter layout."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for cin, cout, kk in _LAYERS:
        blk = np.zeros((cout, _kp(kk)), dtype=np.float32)
        blk[:, :kk] = rng.standard_normal((cout, kk)) * np.sqrt(2.0 / kk)
        out.append(blk.ravel())
    return np.concatenate(out)


def _descs(images, need_centroids: bool):
    torch = _torch()
    d = (ImageDesc * len(images))()
    for i, im in enumerate(images):
        p, n = _points(im["points"])
        w, h = int(im["width"]), int(im["height"])
        if n != w * h:
            raise ValueError(f"image {i}: points must hold width*height pixels")
        c = None
        if need_centroids:
            cen = im["centroids"]
            if tuple(cen.shape) != (3, K):
                raise ValueError(f"image {i}: centroids must be (3, {K})")
            c = _dev(cen, torch.float32, "centroids")
        d[i] = ImageDesc(p, c, None, None, w, h, K if need_centroids else 0, 0)
    return d


def pack(images, out=None):
    """Network input (count, 256, 256, 3) bf16 bits (int16 tensor) of images
    (dicts with points, width, height): area-average resample."""
    torch = _torch()
    d = _descs(images, False)
    n = len(images)
    if out is None:
        out = torch.empty((n, SIDE, SIDE, 3), dtype=torch.int16, device="cuda")
    tmp = torch.empty(max(16 * n, 256), dtype=torch.uint8, device="cuda")
    _check(lib().split_hip_ws_pack(d, n, out.data_ptr(), tmp.data_ptr(), tmp.numel(),
                                   _stream()), "ws_pack")
    return out


def coarse(images, out=None):
    """(count, 3, 256 * 256) float SoA planes of images (dicts with points,
    width, height): the area-average resample the network input uses."""
    torch = _torch()
    d = _descs(images, False)
    n = len(images)
    if out is None:
        out = torch.empty((n, 3, SIDE * SIDE), dtype=torch.float32, device="cuda")
    tmp = torch.empty(max(16 * n, 256), dtype=torch.uint8, device="cuda")
    _check(lib().split_hip_ws_coarse(d, n, out.data_ptr(), tmp.data_ptr(), tmp.numel(),
                                     _stream()), "ws_coarse")
    return out


def refine_on_coarse(images, planes, max_iter: int = 100, threshold: float = 1e-4,
                     labels=None, temp=None):
    """Lloyd iterations on the coarse images (planes, from coarse()) from
    each image's current centroids, written back in place: the clustering
    alone (batched path, connectivity 0).  Returns the coarse iteration
    counts."""
    torch = _torch()
    count = len(images)
    if labels is None:
        labels = torch.empty((count, SIDE * SIDE), dtype=torch.int32, device="cuda")
    cimgs = [dict(points=planes[i], width=SIDE, height=SIDE, centroids=images[i]["centroids"],
                  labels=labels[i]) for i in range(count)]
    if temp is None:
        temp = torch.empty(segment_batch_workspace(cimgs), dtype=torch.uint8, device="cuda")
    res = segment_batch(cimgs, max_iter=max_iter, threshold=threshold, connectivity=0,
                        compress=False, temp=temp)
    return [r.iterations for r in res]


class WarmStartNet:
    """Parameters on the device plus the kernels' workspace."""

    def __init__(self, params: np.ndarray):
        torch = _torch()
        if params.size != int(lib().split_hip_ws_param_count()):
            raise ValueError("parameter vector has the wrong length")
        self.params = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).cuda()
        self._ws = None
        self._ws_batch = 0

    @classmethod
    def load(cls, path: Path = WEIGHTS) -> "WarmStartNet":
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"])

    def save(self, path: Path, meta: dict) -> None:
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, params=self.params.cpu().numpy(),
                            meta=np.array(json.dumps(meta)))

    def workspace(self, batch: int):
        torch = _torch()
        if self._ws is None or self._ws_batch < batch:
            nbytes = int(lib().split_hip_ws_workspace(batch))
            self._ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
            self._ws_batch = batch
        return self._ws

    def forward(self, inputs, index=None, batch=None, out=None):
        torch = _torch()
        B = int(batch if batch is not None else inputs.shape[0])
        ws = self.workspace(B)
        if out is None:
            out = torch.empty((B, OUT), dtype=torch.float32, device="cuda")
        _check(lib().split_hip_ws_forward(self.params.data_ptr(), inputs.data_ptr(),
                                          None if index is None else index.data_ptr(), B,
                                          ws.data_ptr(), ws.numel(), out.data_ptr(),
                                          _stream()), "ws_forward")
        return out

    def backward(self, inputs, targets, grads, loss, index=None, batch=None, kind="mse"):
        B = int(batch if batch is not None else inputs.shape[0])
        ws = self.workspace(B)
        _check(lib().split_hip_ws_backward(self.params.data_ptr(), inputs.data_ptr(),
                                           targets.data_ptr(),
                                           None if index is None else index.data_ptr(), B,
                                           LOSS_KINDS[kind], ws.data_ptr(), ws.numel(),
                                           grads.data_ptr(),
                                           None if loss is None else loss.data_ptr(),
                                           _stream()), "ws_backward")

    def predict(self, images) -> None:
        """Writes the predicted (3, 16) centroids into each image's
        'centroids' tensor."""
        n = len(images)
        ws = self.workspace(n)
        d = _descs(images, True)
        _check(lib().split_hip_ws_predict(self.params.data_ptr(), d, n, ws.data_ptr(),
                                          ws.numel(), _stream()), "ws_predict")

    def predict_refined(self, images, max_iter: int = 100, threshold: float = 1e-4):
        """The shipped warm start: predict(), then Lloyd iterations from the
        prediction on each image's 256^2 coarse resample (refine_on_coarse)
        until shift < threshold or max_iter.  Returns the coarse iteration
        counts."""
        self.predict(images)
        return refine_on_coarse(images, coarse(images), max_iter=max_iter,
                                threshold=threshold)


def adam(params, grads, m, v, step: int, lr: float, b1=0.9, b2=0.999, eps=1e-8) -> None:
    _check(lib().split_hip_ws_adam(params.data_ptr(), grads.data_ptr(), m.data_ptr(),
                                   v.data_ptr(), int(step), float(lr), float(b1), float(b2),
                                   float(eps), _stream()), "ws_adam")


def targets(centroids_flat, count: int, out=None):
    """(count, 48) luminance-sorted targets from (count, 3, 16) centroids."""
    torch = _torch()
    if out is None:
        out = torch.empty((count, OUT), dtype=torch.float32, device="cuda")
    _check(lib().split_hip_ws_targets(centroids_flat.data_ptr(), count, out.data_ptr(),
                                      _stream()), "ws_targets")
    return out


# ---------------------------------------------------------------- data ---

def training_specs(count: int, seed: int, side: int = SIDE):
    """Config-2-style images (Voronoi x Lambertian shading, noise 0.01) of
    side x side with 20..300 sites, drawn from a seeded PCG64 stream."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        sites = int(rng.integers(20, 301))
        out.append((W.ImageSpec(side, side, seed=int(rng.integers(0, 2**63)), sites=sites,
                                shading=True, noise_sigma=0.01),
                    int(rng.integers(0, 2**63))))
    return out


def make_dataset(count: int, seed: int, chunk: int = 1024, log=print):
    """Inputs (count, 256, 256, 3) bf16 and targets (count, 48) on the device:
    each image generated on the device, clustered by plain k-means (batched
    path), its converged centroids sorted by luminance."""
    torch = _torch()
    specs = training_specs(count, seed)
    inputs = torch.empty((count, SIDE, SIDE, 3), dtype=torch.int16, device="cuda")
    tgt = torch.empty((count, OUT), dtype=torch.float32, device="cuda")
    n = SIDE * SIDE
    planes = torch.empty((chunk, 3, n), dtype=torch.float32, device="cuda")
    cent = torch.empty((chunk, 3, K), dtype=torch.float32, device="cuda")
    labels = torch.empty((chunk, n), dtype=torch.int32, device="cuda")
    segs = torch.empty((chunk, n), dtype=torch.int32, device="cuda")
    from . import gen_images
    iters = []
    temp = None
    for c0 in range(0, count, chunk):
        c1 = min(count, c0 + chunk)
        sub = specs[c0:c1]
        gen_images([s for s, _ in sub], [planes[i] for i in range(c1 - c0)])
        imgs = []
        for i, (s, init_seed) in enumerate(sub):
            uniform_random_initialize(planes[i], cent[i], init_seed)
            imgs.append(dict(points=planes[i], width=SIDE, height=SIDE, centroids=cent[i],
                             labels=labels[i], segments=segs[i]))
        if temp is None or c1 - c0 != chunk:
            temp = torch.empty(segment_batch_workspace(imgs), dtype=torch.uint8,
                               device="cuda")
        res = segment_batch(imgs, max_iter=100, threshold=1e-4, connectivity=8,
                            compress=False, temp=temp)
        iters += [r.iterations for r in res]
        targets(cent[:c1 - c0], c1 - c0, out=tgt[c0:c1])
        pack(imgs, out=inputs[c0:c1])
        if log:
            log(f"dataset: {c1}/{count} images, mean iterations {np.mean(iters):.1f}")
    torch.cuda.synchronize()
    return inputs, tgt, iters


def train(images: int = 50000, epochs: int = 12, batch: int = 256, lr: float = 2e-3,
          seed: int = 0x57A4_0001, holdout: int = 1024, loss_kind: str = "chamfer",
          log=print):
    torch = _torch()
    t0 = time.time()
    inputs, tgt, iters = make_dataset(images + holdout, seed, log=log)
    log(f"dataset ready in {time.time() - t0:.1f} s")
    net = WarmStartNet(init_params(seed ^ 0x1234))
    P = param_count()
    grads = torch.zeros(P, dtype=torch.float32, device="cuda")
    m = torch.zeros_like(grads)
    v = torch.zeros_like(grads)
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    rng = np.random.Generator(np.random.PCG64(seed ^ 0x5678))
    steps_per_epoch = images // batch
    total = steps_per_epoch * epochs
    step = 0
    hist = []
    hidx = torch.arange(images, images + holdout, dtype=torch.int32, device="cuda")

    def holdout_loss():
        tot = 0.0
        for h0 in range(0, holdout, batch):
            hb = min(batch, holdout - h0)
            net.forward(inputs, index=hidx[h0:h0 + hb], batch=hb)
            net.backward(inputs, tgt, grads, loss, index=hidx[h0:h0 + hb], batch=hb,
                         kind=loss_kind)
            tot += float(loss.item()) * hb
        return tot / holdout

    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(images).astype(np.int32)).cuda()
        run = 0.0
        for s in range(steps_per_epoch):
            step += 1
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * (step - 1) / total))
            idx = perm[s * batch:(s + 1) * batch]
            net.forward(inputs, index=idx, batch=batch)
            net.backward(inputs, tgt, grads, loss, index=idx, batch=batch, kind=loss_kind)
            adam(net.params, grads, m, v, step, lr_t)
            if s % 50 == 0:
                run = float(loss.item())
        hl = holdout_loss()
        hist.append({"epoch": ep + 1, "train_loss": run, "holdout_loss": hl})
        log(f"epoch {ep + 1}/{epochs}: train loss {run:.5f}, holdout loss {hl:.5f}, "
            f"{time.time() - t0:.1f} s")
    meta = {"images": images, "holdout": holdout, "epochs": epochs, "batch": batch, "lr": lr,
            "seed": seed, "loss": loss_kind, "history": hist, "target_iterations_mean": float(np.mean(iters))}
    return net, meta


def eval_specs(count: int = 256, side: int = 1024, seed: int = 0xE7A1_0004):
    """Evaluation images: config-2-style side x side images (20..300 sites),
    seeds disjoint from the training stream."""
    return training_specs(count, seed, side)


def evaluate(net: "WarmStartNet", count: int = 256, side: int = 1024, reps: int = 1,
             separate: bool = False, log=print) -> dict:
    """Plain (seeded-sample init) against warm-started batched segmentation
    of the evaluation images: Lloyd iterations to shift < 1e-4 (at most 100),
    final inertia, and time per run (init + segmentation).  Warm starts:
    'warm' = the network's prediction; 'warm_coarse' = the prediction refined
    by Lloyd iterations on the 256^2 coarse image (coarse() +
    refine_on_coarse(), shift < 1e-4); 'plain_coarse' (the ablation) = the
    seeded sample refined the same way.  With `separate`, also the plain
    method as one segment() call per image (the unbatched GPU path), whose
    results equal the batched plain run."""
    torch = _torch()
    from . import gen_images, inertia, segment, segment_workspace
    specs = eval_specs(count, side)
    n = side * side
    planes = torch.empty((count, 3, n), dtype=torch.float32, device="cuda")
    gen_images([s for s, _ in specs], [planes[i] for i in range(count)])
    cent = torch.empty((count, 3, K), dtype=torch.float32, device="cuda")
    labels = torch.empty((count, n), dtype=torch.int32, device="cuda")
    segs = torch.empty((count, n), dtype=torch.int32, device="cuda")
    imgs = [dict(points=planes[i], width=side, height=side, centroids=cent[i],
                 labels=labels[i], segments=segs[i]) for i in range(count)]
    temp = torch.empty(segment_batch_workspace(imgs), dtype=torch.uint8, device="cuda")
    itemp = torch.empty(int(lib().split_hip_inertia_workspace()), dtype=torch.uint8,
                        device="cuda")
    cplanes = torch.empty((count, 3, SIDE * SIDE), dtype=torch.float32, device="cuda")
    clabels = torch.empty((count, SIDE * SIDE), dtype=torch.int32, device="cuda")
    ctemp = torch.empty(segment_batch_workspace(
        [dict(points=cplanes[i], width=SIDE, height=SIDE, centroids=cent[i],
              labels=clabels[i]) for i in range(count)]), dtype=torch.uint8, device="cuda")
    net.workspace(count)
    coarse_iters = {}

    def plain_init():
        for i, (_, init_seed) in enumerate(specs):
            uniform_random_initialize(planes[i], cent[i], init_seed)

    def warm_init():
        net.predict(imgs)

    def refined(init, name):
        def run():
            init()
            coarse(imgs, out=cplanes)
            coarse_iters[name] = refine_on_coarse(imgs, cplanes, labels=clabels, temp=ctemp)
        return run

    out = {}
    for name, init in (("plain", plain_init), ("warm", warm_init),
                       ("warm_coarse", refined(warm_init, "warm_coarse")),
                       ("plain_coarse", refined(plain_init, "plain_coarse"))):
        init()
        segment_batch(imgs, max_iter=100, threshold=1e-4, connectivity=8, temp=temp)
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            init()
            res = segment_batch(imgs, max_iter=100, threshold=1e-4, connectivity=8,
                                temp=temp)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        its = np.array([r.iterations for r in res])
        ine = np.array([inertia(planes[i], cent[i], temp=itemp) for i in range(count)])
        out[name] = {"seconds": best, "iterations": its, "inertia": ine,
                     "segments": np.array([r.num_segments for r in res])}
        if name in coarse_iters:
            out[name]["coarse_iterations"] = np.array(coarse_iters[name])
        if log:
            extra = (f", coarse iterations mean {np.mean(coarse_iters[name]):.2f}"
                     if name in coarse_iters else "")
            log(f"eval {name}: {best * 1e3:.1f} ms, iterations mean {its.mean():.2f} "
                f"(median {np.median(its):.0f}, hit cap {(its >= 100).sum()}){extra}, "
                f"inertia sum {ine.sum():.2f}")
    if separate:
        stemp = torch.empty(segment_workspace(side, side, K), dtype=torch.uint8,
                            device="cuda")

        def run_separate():
            res = []
            for i, (_, init_seed) in enumerate(specs):
                uniform_random_initialize(planes[i], cent[i], init_seed)
                res.append(segment(planes[i], side, side, cent[i], labels[i], segs[i],
                                   max_iter=100, threshold=1e-4, connectivity=8,
                                   temp=stemp))
            return res

        run_separate()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = run_separate()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        same = all(a.iterations == b for a, b in zip(res, out["plain"]["iterations"]))
        out["separate"] = {"seconds": dt, "same_iterations": same}
        if log:
            log(f"eval separate plain calls: {dt * 1e3:.1f} ms (same iterations: {same})")
    P = out["plain"]
    summ = {"images": count, "side": side, "pixels": count * n,
            "plain_ms": P["seconds"] * 1e3,
            "plain_mpix_s": count * n / P["seconds"] / 1e6,
            "plain_iters_mean": float(P["iterations"].mean()),
            "plain_cap_hits": int((P["iterations"] >= 100).sum())}
    for name in ("warm", "warm_coarse", "plain_coarse"):
        o = out[name]
        ratio = o["inertia"] / P["inertia"]
        summ.update({
            f"{name}_ms": o["seconds"] * 1e3,
            f"{name}_mpix_s": count * n / o["seconds"] / 1e6,
            f"{name}_iters_mean": float(o["iterations"].mean()),
            f"{name}_cap_hits": int((o["iterations"] >= 100).sum()),
            f"{name}_inertia_ratio_total": float(o["inertia"].sum() / P["inertia"].sum()),
            f"{name}_inertia_ratio_median": float(np.median(ratio)),
            f"{name}_images_worse_0p5pct": int((ratio > 1.005).sum()),
        })
        if "coarse_iterations" in o:
            ci = float(o["coarse_iterations"].mean())
            summ[f"{name}_coarse_stage_iters_mean"] = ci
            # coarse iterations at their pixel share of a full-size one
            summ[f"{name}_pixel_iters_mean"] = float(o["iterations"].mean()) + ci * SIDE * SIDE / n
    # earlier names of the network-only figures
    summ.update({"warm_mpix_s": summ["warm_mpix_s"], "speedup": P["seconds"] / out["warm"]["seconds"],
                 "inertia_ratio_total": summ["warm_inertia_ratio_total"]})
    if separate:
        sep = out["separate"]["seconds"]
        summ.update({
            "separate_plain_ms": sep * 1e3,
            "separate_plain_mpix_s": count * n / sep / 1e6,
            "speedup_batched_warm_vs_separate_plain": sep / out["warm"]["seconds"],
            "speedup_batched_warm_coarse_vs_separate_plain": sep / out["warm_coarse"]["seconds"],
            "separate_same_iterations": out["separate"]["same_iterations"],
        })
    out["summary"] = summ
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--images", type=int, default=50000)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--loss", choices=sorted(LOSS_KINDS), default="chamfer")
    t.add_argument("--out", type=Path, default=WEIGHTS)
    t.add_argument("--eval", type=int, default=256, help="evaluation images (0: none)")
    e = sub.add_parser("eval")
    e.add_argument("--weights", type=Path, default=WEIGHTS)
    e.add_argument("--images", type=int, default=256)
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.cmd == "train":
        net, meta = train(args.images, args.epochs, args.batch, args.lr, loss_kind=args.loss,
                          log=log)
        if args.eval:
            meta["eval"] = evaluate(net, args.eval, log=log)["summary"]
            log(json.dumps(meta["eval"]))
        net.save(args.out, meta)
        log(f"saved {args.out}")
    else:
        net = WarmStartNet.load(args.weights)
        print(json.dumps(evaluate(net, args.images, separate=True, log=log)["summary"]))


if __name__ == "__main__":
    main()
