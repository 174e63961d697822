"""Exact centroid sums at the feature scales the reference clusters.

The reference's sample clusters (beta, a, b) features: Lab of the image with
the beta feature (values -1e5, 0, +1e5) written over L
(samples/src/main.cu:216-245, beta_feature.cu:48-78), a and b up to about
+-128.  The tiled, batched and segment-mean paths sum quantised integers:
each channel c has its own power-of-two scale s_c with |rint(v s_c)| <=
2^23 - 1, so a mean moves by at most half a quantum, 1 / (2 s_c) <=
max|v_c| 2^-23, whatever the other channels hold.

Bars (per channel c, against the FP64 mean of the GPU's own labels):
  |c_gpu - c_fp64| <= 1 / (2 s_c) + half a float32 ulp of c_fp64
(the quantisation bound plus the rounding of the float result), and that
bound is <= 1e-5 * max(1, max|v_c|) — 1e-5 absolute for colours in [0, 1],
1e-5 of the channel's range for Lab and beta.  Labels: the labels after I
iterations equal a dense assignment (label_image) against the centroids
after I - 1 iterations.  Segments: the CPU union-find of the labels.
"""
import numpy as np
import pytest

import oracle
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

CENT_REL = 1e-5


def _scale(m):
    """The documented per-channel scale for a channel whose largest |v| is m."""
    m = np.float32(m)
    e = int(np.frexp(m)[1]) if m > 0 else 0
    e = max(e, -100)
    s = np.ldexp(1.0, 23 - e)
    if float(m) * s > 8388607.0:
        s *= 0.5
    return s


def check_means(pts, labels, cent, what):
    """cent (3, k) against the FP64 means of pts over labels (see the module
    bars); returns the number of non-empty clusters."""
    k = cent.shape[1]
    lab = labels.astype(np.int64)
    cnt = np.bincount(lab, minlength=k)
    live = cnt > 0
    got = cent.astype(np.float64)
    for c in range(3):
        s = np.bincount(lab, weights=pts[c].astype(np.float64), minlength=k)
        want = s / np.maximum(cnt, 1)
        m = float(np.abs(pts[c]).max())
        quant = 0.5 / _scale(m)
        tol = quant + 0.5 * np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
        assert quant <= CENT_REL * max(1.0, m), (what, c, quant, m)
        err = np.abs(got[c] - want)
        bad = live & (err > tol + 1e-12)
        assert not bad.any(), (what, c, float(err[live].max()), float(tol[live].min()))
        # the bound holds with room: the mean of rounding errors is small
        assert np.all(err[live] <= CENT_REL * max(1.0, m)), (what, c)
    return int(live.sum())


def features(torch, img, kind):
    """(3, N) feature planes of an RGB image on the device."""
    P = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    F = torch.empty_like(P)
    if kind == "rgb":
        F.copy_(P)
    elif kind == "lab":
        sa.convert_color_space(P, F, "rgb_to_lab")
    elif kind == "beta_ab":
        sa.lab_beta_feature(P, F, None, mu=1e5)
    else:
        raise ValueError(kind)
    torch.cuda.synchronize()
    return F


def seeded(torch, F, k, seed):
    C = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.uniform_random_initialize(F, C, seed)
    return C


@pytest.fixture(scope="module")
def c1(torch_cuda):
    return W.make_image(W.PLAIN_512.image)


@pytest.mark.parametrize("kind,k", [("lab", 8), ("beta_ab", 8), ("beta_ab", 5), ("lab", 64)])
def test_segment_feature_scales(torch_cuda, c1, kind, k):
    """segment (tiled bounded Lloyd + CC + compress) on Lab and (beta, a, b)
    features of the 512x512 image, 20 iterations."""
    torch = torch_cuda
    iters = 20
    say(f"scale: segment {kind} k={k} ({iters - 1} and {iters} iterations)")
    F = features(torch, c1, kind)
    init = seeded(torch, F, k, 0x5CA1E + k)
    runs = {}
    for it in (iters - 1, iters):
        C = init.clone()
        L = torch.empty(F.shape[1], dtype=torch.int32, device="cuda")
        S = torch.empty(F.shape[1], dtype=torch.int32, device="cuda")
        r = sa.segment(F, 512, 512, C, L, S, max_iter=it, threshold=-1.0, connectivity=4)
        assert r.iterations == it
        runs[it] = (C, L, S, r)
    C19, _, _, _ = runs[iters - 1]
    C20, L20, S20, r20 = runs[iters]
    dense = torch.empty_like(L20)
    sa.label_points(C19, F, dense, width=512, height=512)
    assert torch.equal(dense, L20)
    pts = F.cpu().numpy()
    lab = L20.cpu().numpy()
    assert check_means(pts, lab, C20.cpu().numpy(), f"segment {kind}") >= 2
    rep, _ = oracle.cc_uf(lab.reshape(512, 512), 4)
    want, cnt = oracle.compress(rep)
    assert r20.num_segments == cnt
    assert np.array_equal(S20.cpu().numpy(), want)


@pytest.mark.parametrize("kind", ["lab", "beta_ab"])
def test_cluster_image_feature_scales(torch_cuda, c1, kind):
    """cluster with an image width (tiled bounded path), k = 16, against the
    FP64 means of its labels; labels equal a dense assignment against the
    previous iteration's centroids."""
    torch = torch_cuda
    k, iters = 16, 12
    say(f"scale: cluster_image {kind} k={k}")
    F = features(torch, c1, kind)
    init = seeded(torch, F, k, 0xC1A55 + k)
    out = {}
    for it in (iters - 1, iters):
        C = init.clone()
        L = torch.empty(F.shape[1], dtype=torch.int32, device="cuda")
        n_it, _ = sa.cluster(F, C, L, it, -1.0, width=512, height=512)
        assert n_it == it
        out[it] = (C, L)
    dense = torch.empty_like(out[iters][1])
    sa.label_points(out[iters - 1][0], F, dense, width=512, height=512)
    assert torch.equal(dense, out[iters][1])
    check_means(F.cpu().numpy(), out[iters][1].cpu().numpy(), out[iters][0].cpu().numpy(),
                f"cluster_image {kind}")


@pytest.mark.parametrize("kind", ["lab", "beta_ab"])
def test_segment_means_feature_scales(torch_cuda, c1, kind):
    """calculate_centroids with k = number of segments (> 1024: the segment
    mean kernels) on feature planes."""
    torch = torch_cuda
    say(f"scale: segment means {kind}")
    F = features(torch, c1, kind)
    C = seeded(torch, F, 8, 0x5E6 + 1)
    L = torch.empty(F.shape[1], dtype=torch.int32, device="cuda")
    S = torch.empty(F.shape[1], dtype=torch.int32, device="cuda")
    r = sa.segment(F, 512, 512, C, L, S, max_iter=20, threshold=-1.0, connectivity=4)
    nseg = r.num_segments
    assert nseg > 1024, nseg
    M = torch.empty((3, nseg), dtype=torch.float32, device="cuda")
    sa.calculate_centroids(S, F, M)
    assert check_means(F.cpu().numpy(), S.cpu().numpy(), M.cpu().numpy(),
                       f"segment means {kind}") == nseg


def test_segment_batch_feature_scales(torch_cuda):
    """segment_batch over images of RGB, Lab and (beta, a, b) features of
    uneven sizes and k: per image, centroids against the FP64 means of its
    labels, labels against a dense assignment with the centroids of an
    (I - 1)-iteration batch, segments against the CPU union-find."""
    torch = torch_cuda
    shapes = [(300, 200, 8, "lab"), (257, 129, 5, "beta_ab"), (512, 384, 16, "beta_ab"),
              (333, 517, 32, "lab"), (128, 128, 64, "rgb"), (1023, 211, 12, "beta_ab")]
    iters = 15
    say(f"scale: segment_batch of {len(shapes)} feature images, {iters - 1} and {iters} its")
    feats, inits = [], []
    for i, (w, h, k, kind) in enumerate(shapes):
        spec = W.ImageSpec(w, h, seed=0x5CA1_E000 + i, sites=40 + 10 * i, shading=True,
                           noise_sigma=0.01, base_cell=max(w, h) / 4.0)
        F = features(torch, W.make_image(spec), kind)
        feats.append(F)
        inits.append(seeded(torch, F, k, 0x1000 + i))
    runs = {}
    for it in (iters - 1, iters):
        imgs = []
        for (w, h, k, kind), F, C0 in zip(shapes, feats, inits):
            imgs.append(dict(points=F, width=w, height=h, centroids=C0.clone(),
                             labels=torch.full((w * h,), -7, dtype=torch.int32, device="cuda"),
                             segments=torch.full((w * h,), -7, dtype=torch.int32,
                                                 device="cuda")))
        res = sa.segment_batch(imgs, max_iter=it, threshold=-1.0, connectivity=8)
        torch.cuda.synchronize()
        runs[it] = (imgs, res)
    for i, ((w, h, k, kind), F) in enumerate(zip(shapes, feats)):
        tag = f"image {i} {w}x{h} k={k} {kind}"
        prev, last = runs[iters - 1][0][i], runs[iters][0][i]
        r = runs[iters][1][i]
        assert r.iterations == iters, tag
        dense = torch.empty_like(last["labels"])
        sa.label_points(prev["centroids"], F, dense, width=w, height=h)
        assert torch.equal(dense, last["labels"]), tag
        lab = last["labels"].cpu().numpy()
        check_means(F.cpu().numpy(), lab, last["centroids"].cpu().numpy(), tag)
        rep, _ = oracle.cc_uf(lab.reshape(h, w), 8)
        want, cnt = oracle.compress(rep)
        assert r.num_segments == cnt, tag
        assert np.array_equal(last["segments"].cpu().numpy(), want), tag
