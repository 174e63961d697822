"""Batched segmentation (split_hip_segment_batch) through the C ABI.

Bar: every image's outputs — final centroids, cluster labels, segment
labels, iteration count, last shift and segment count — are bit-identical to
split_hip_segment of that image alone (whose parity against the oracle is
covered by test_gpu_parity), on images of uneven size, alignment and k, with
and without graph replay.  The segment labels of one image are also checked
directly against the oracle's union-find.
"""
import numpy as np
import pytest

import oracle
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _specs():
    # (width, height, k, sites): ragged edges, widths not multiple of 4,
    # thin strips, one tile, k = 1 and the batch maximum.
    shapes = [(37, 53, 4, 20), (130, 97, 8, 40), (300, 200, 16, 60), (1, 500, 2, 5),
              (16, 16, 1, 3), (257, 129, 32, 80), (512, 384, 64, 150), (500, 3, 8, 4),
              (96, 64, 64, 30), (1023, 211, 16, 120), (64, 64, 5, 10), (333, 517, 12, 90)]
    out = []
    for i, (w, h, k, sites) in enumerate(shapes):
        spec = W.ImageSpec(w, h, seed=0xBA7C_0000 + i, sites=sites, shading=True,
                           noise_sigma=0.01, base_cell=max(w, h) / 4.0)
        out.append((spec, k, 0x1417_0100 + i))
    return out


@pytest.fixture(scope="module")
def batch_images(torch_cuda):
    torch = torch_cuda
    imgs = []
    for spec, k, seed in _specs():
        img = W.make_image(spec)
        P = torch.from_numpy(img).cuda()
        C0 = torch.empty((3, k), dtype=torch.float32, device="cuda")
        sa.uniform_random_initialize(P, C0, seed)
        imgs.append(dict(points=P, width=spec.width, height=spec.height, init=C0.clone()))
    torch.cuda.synchronize()
    return imgs


def _fresh(torch, imgs):
    out = []
    for im in imgs:
        n = im["width"] * im["height"]
        out.append(dict(points=im["points"], width=im["width"], height=im["height"],
                        centroids=im["init"].clone(),
                        labels=torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                        segments=torch.full((n,), -7, dtype=torch.int32, device="cuda")))
    return out


def _single(torch, imgs, max_iter, threshold, conn, compress):
    outs = _fresh(torch, imgs)
    res = []
    for im in outs:
        res.append(sa.segment(im["points"], im["width"], im["height"], im["centroids"],
                              im["labels"], im["segments"], max_iter=max_iter,
                              threshold=threshold, connectivity=conn, compress=compress))
    torch.cuda.synchronize()
    return outs, res


def _assert_same(torch, got, got_res, want, want_res, compress):
    for i, (g, gr, w, wr) in enumerate(zip(got, got_res, want, want_res)):
        tag = f"image {i} ({w['width']}x{w['height']}, k={w['centroids'].shape[1]})"
        assert gr.iterations == wr.iterations, tag
        assert gr.shift == wr.shift, tag
        if compress:
            assert gr.num_segments == wr.num_segments, tag
        else:
            assert gr.num_segments == -1, tag
        assert torch.equal(g["centroids"], w["centroids"]), tag
        assert torch.equal(g["labels"], w["labels"]), tag
        assert torch.equal(g["segments"], w["segments"]), tag


@pytest.mark.parametrize("max_iter,threshold,conn,compress,use_graph", [
    (20, -1.0, 4, True, False),
    (20, -1.0, 8, True, True),
    (60, 1e-4, 8, True, True),
    (60, 1e-3, 4, False, True),
    (0, 1e-4, 8, True, True),
])
def test_batch_equals_single(torch_cuda, batch_images, max_iter, threshold, conn, compress,
                             use_graph):
    torch = torch_cuda
    say(f"batch vs single: iters={max_iter} thr={threshold} conn={conn} "
        f"compress={compress} graph={use_graph}")
    want, want_res = _single(torch, batch_images, max_iter, threshold, conn, compress)
    got = _fresh(torch, batch_images)
    got_res = sa.segment_batch(got, max_iter=max_iter, threshold=threshold,
                               connectivity=conn, compress=compress, use_graph=use_graph)
    torch.cuda.synchronize()
    _assert_same(torch, got, got_res, want, want_res, compress)
    if threshold > 0 and max_iter > 0:
        iters = {r.iterations for r in got_res}
        assert len(iters) > 1, "images should stop at different iterations"


def test_batch_graph_replay(torch_cuda, batch_images):
    """The captured graph replays with new buffers in the descriptor table."""
    torch = torch_cuda
    want, want_res = _single(torch, batch_images, 40, 1e-4, 8, True)
    temp = torch.empty(sa.segment_batch_workspace(_fresh(torch, batch_images)),
                       dtype=torch.uint8, device="cuda")
    for rep in range(3):
        say(f"batch graph replay {rep}")
        got = _fresh(torch, batch_images)  # new output buffers each replay
        got_res = sa.segment_batch(got, max_iter=40, threshold=1e-4, connectivity=8,
                                   compress=True, use_graph=True, temp=temp)
        torch.cuda.synchronize()
        _assert_same(torch, got, got_res, want, want_res, True)
    # a smaller batch in the same temp
    say("batch graph: sub-batch")
    sub = batch_images[3:7]
    want, want_res = _single(torch, sub, 40, 1e-4, 8, True)
    got = _fresh(torch, sub)
    got_res = sa.segment_batch(got, max_iter=40, threshold=1e-4, connectivity=8,
                               compress=True, use_graph=True, temp=temp)
    torch.cuda.synchronize()
    _assert_same(torch, got, got_res, want, want_res, True)


def test_batch_segments_vs_oracle(torch_cuda, batch_images):
    torch = torch_cuda
    say("batch segments vs oracle union-find")
    got = _fresh(torch, batch_images)
    res = sa.segment_batch(got, max_iter=15, threshold=-1.0, connectivity=8)
    torch.cuda.synchronize()
    for g, r in zip(got, res):
        lab = g["labels"].cpu().numpy().reshape(g["height"], g["width"])
        rep, ncomp = oracle.cc_uf(lab, 8)
        want, cnt = oracle.compress(rep)
        assert r.num_segments == cnt == ncomp
        assert np.array_equal(g["segments"].cpu().numpy(), want)


def test_batch_errors(torch_cuda, batch_images):
    torch = torch_cuda
    say("batch argument errors")
    im = _fresh(torch, batch_images[:1])[0]
    big = dict(im, centroids=torch.zeros((3, 65), dtype=torch.float32, device="cuda"))
    with pytest.raises(sa.SplitError) as e:
        sa.segment_batch([big], max_iter=5)
    assert e.value.status == 5  # UNSUPPORTED
    with pytest.raises(sa.SplitError) as e:
        sa.segment_batch([im], max_iter=5, connectivity=6)
    assert e.value.status == 1
    small = torch.empty(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(sa.SplitError) as e:
        sa.segment_batch([im], max_iter=5, temp=small)
    assert e.value.status == 2  # TEMP


def test_batch_graph_key_cc_tiles(torch_cuda):
    """Two one-image batches with the same k and padded pixel count (2304)
    (and, with 16x16 Lloyd tiles, the same tile count, 9) but different
    counts of 64x64 CC tiles (48x48: 1, 144x16: 3) share one temp: the
    second call must capture its own graph, and its segments equal the CPU
    union-find."""
    torch = torch_cuda
    imgs = []
    for i, (w, h) in enumerate([(48, 48), (144, 16)]):
        spec = W.ImageSpec(w, h, seed=0x6EA7_0000 + i, sites=12, shading=True,
                           noise_sigma=0.01, base_cell=max(w, h) / 4.0)
        P = torch.from_numpy(W.make_image(spec)).cuda()
        C0 = torch.empty((3, 8), dtype=torch.float32, device="cuda")
        sa.uniform_random_initialize(P, C0, 0x6EA7 + i)
        imgs.append(dict(points=P, width=w, height=h, init=C0))
    sizes = [sa.segment_batch_workspace(_fresh(torch, [im])) for im in imgs]
    temp = torch.empty(max(sizes), dtype=torch.uint8, device="cuda")
    for rep in range(2):
        for im in imgs:
            say(f"batch graph key: {im['width']}x{im['height']} pass {rep}")
            got = _fresh(torch, [im])
            res = sa.segment_batch(got, max_iter=10, threshold=-1.0, connectivity=8,
                                   compress=True, use_graph=True, temp=temp)
            torch.cuda.synchronize()
            g = got[0]
            lab = g["labels"].cpu().numpy().reshape(g["height"], g["width"])
            rep_, ncomp = oracle.cc_uf(lab, 8)
            want, cnt = oracle.compress(rep_)
            assert res[0].num_segments == cnt == ncomp
            assert np.array_equal(g["segments"].cpu().numpy(), want)
