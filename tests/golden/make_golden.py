"""Writes tests/golden/*.npz: small seeded inputs of the segmentation path
with the oracle's outputs (FP64 labels and centroids after a fixed number of
Lloyd iterations, reference-restated float labels, union-find components at
4 and 8 connectivity, dense ranks).  The oracle functions behind them are
pinned in tests/test_oracle.py; these files freeze their values so that a
change in the generator or the oracle is noticed, and give the GPU tests
fixed vectors.

    python tests/golden/make_golden.py
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import oracle  # noqa: E402
from split_amd import workloads as W  # noqa: E402

CASES = [
    # name, width, height, sites, shading, sigma, k, iters, init seed
    ("voronoi_16x16", 16, 16, 6, False, 0.02, 4, 5, 11),
    ("voronoi_64x64", 64, 64, 12, False, 0.02, 8, 20, 12),
    ("shaded_64x48", 64, 48, 10, True, 0.01, 8, 10, 13),
]


def make(name, width, height, sites, shading, sigma, k, iters, seed):
    spec = W.ImageSpec(width, height, seed=0x60_1DE + seed, sites=sites, shading=shading,
                       noise_sigma=sigma)
    img = W.make_image(spec)
    idx = oracle.sample_indices(seed, width * height, k)
    init = np.ascontiguousarray(img[:, idx])
    cent, lab = oracle.lloyd_f64(img, init, iters)
    ref_lab = oracle.ref_label_points(init, img)
    cc4, n4 = oracle.cc_uf(lab.reshape(height, width), 4)
    cc8, n8 = oracle.cc_uf(lab.reshape(height, width), 8)
    rank8, _ = oracle.compress(cc8)
    np.savez_compressed(
        HERE / f"{name}.npz", image=img, width=width, height=height, sites=sites,
        shading=int(shading), sigma=sigma, gen_seed=spec.seed, k=k, iters=iters,
        init_seed=seed, init_index=idx, init=init, centroids=cent, labels=lab,
        ref_labels_init=ref_lab, cc4=cc4, cc8=cc8, ncc4=n4, ncc8=n8, rank8=rank8)
    print(f"{name}: {width}x{height} k={k} iters={iters} cc4={n4} cc8={n8}")


if __name__ == "__main__":
    for c in CASES:
        make(*c)
