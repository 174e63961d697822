"""Full-size cases — the benchmark shapes: 4096² with k = 32 for 100
iterations, 8192² with k = 256, and the 8192² spiral/checkerboard CC image —
checked through properties that hold at any size (the CPU oracle cannot redo
whole runs of this size in test time):

* the bounded iterations' labels, which skip every pixel whose label is
  provably unchanged, equal a dense assignment of every pixel against the
  previous iteration's centroids;
* the centroids are, within 1e-5, the FP64 means of the GPU's own labels;
* the components equal a CPU union-find of the GPU's labels (largest linear
  index per component), and the compressed segments their dense rank.

The images come from the device generator (the host generator's images up
to the last bit of the noise terms); the properties hold for any input."""
import numpy as np
import pytest

import oracle
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

CENT_TOL = 1e-5


def _image(torch, w):
    P = torch.empty((3, w.n), dtype=torch.float32, device="cuda")
    sa.gen_images([w.image], [P])
    return P


def _fp64_means(pts, labels, k):
    cnt = np.bincount(labels, minlength=k).astype(np.float64)
    out = np.empty((3, k))
    for c in range(3):
        s = np.bincount(labels, weights=pts[c].astype(np.float64), minlength=k)
        out[c] = s / np.maximum(cnt, 1.0)
    return out, cnt


@pytest.mark.parametrize("name,iters", [("shaded_4096", 100), ("stress_8192", 24)])
def test_fullsize_bounded_lloyd(torch_cuda, name, iters):
    torch = torch_cuda
    w = W.WORKLOADS[name]
    wd, ht = w.image.width, w.image.height
    say(f"full size {name} k={w.k}: {iters - 1} and {iters} iterations")
    P = _image(torch, w)
    init = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    sa.uniform_random_initialize(P, init, w.init_seed)
    runs = {}
    for n_it in (iters - 1, iters):
        C = init.clone()
        L = torch.empty(w.n, dtype=torch.int32, device="cuda")
        it, _ = sa.cluster(P, C, L, n_it, -1.0, width=wd, height=ht)
        assert it == n_it
        runs[n_it] = (C, L)
    c_prev, _ = runs[iters - 1]
    c_last, l_last = runs[iters]
    dense = torch.empty(w.n, dtype=torch.int32, device="cuda")
    sa.label_points(c_prev, P, dense, width=wd, height=ht)
    assert torch.equal(dense, l_last)
    lab = l_last.cpu().numpy()
    means, cnt = _fp64_means(P.cpu().numpy(), lab, w.k)
    live = cnt > 0
    assert live.sum() >= w.k // 2
    got = c_last.cpu().numpy().astype(np.float64)
    assert np.abs(got[:, live] - means[:, live]).max() <= CENT_TOL


def test_fullsize_segment_shaded_4096(torch_cuda):
    """The default benchmark step end to end: segments = the CPU union-find
    of the GPU's final labels, compressed."""
    torch = torch_cuda
    w = W.SHADED_4096
    say("full size shaded_4096 segment (100 iterations, 8-connectivity)")
    P = _image(torch, w)
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    sa.uniform_random_initialize(P, C, w.init_seed)
    L = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    r = sa.segment(P, 4096, 4096, C, L, S, max_iter=w.max_iter, threshold=w.threshold,
                   connectivity=w.connectivity)
    assert 1 <= r.iterations <= w.max_iter
    rep, ncomp = oracle.cc_uf(L.cpu().numpy().reshape(4096, 4096), w.connectivity)
    want, cnt = oracle.compress(rep)
    assert r.num_segments == cnt == ncomp
    assert np.array_equal(S.cpu().numpy(), want)


def test_fullsize_cc_stress_8192(torch_cuda):
    """Nested 1-pixel spirals and a checkerboard band at 8192², 8-connectivity:
    components = the CPU union-find's, and compress counts them."""
    torch = torch_cuda
    say("full size cc_stress 8192 (8-connectivity)")
    lab = W.make_cc_stress(8192, 8192)
    L = torch.from_numpy(lab).cuda()
    out = torch.empty(lab.size, dtype=torch.int32, device="cuda")
    sa.connected_components(L, out, connectivity=8)
    rep, ncomp = oracle.cc_uf(lab, 8)
    assert np.array_equal(out.cpu().numpy(), rep)
    assert sa.compress_labels(out) == ncomp


def test_fullsize_batch_256(torch_cuda):
    """The batched configuration at its size: all 256 images of
    batch_workloads() (sides 128..2048, k in {4..64}, ~3e8 pixels) in one
    segment_batch call (shift < 1e-4 or 100 iterations, 8-connectivity).
    Per image, against the oracle and properties, not against the single-
    image path: centroids within 1e-5 of the FP64 means of the image's own
    labels; labels equal a dense assignment against the centroids after
    I - 1 iterations (I = the image's iteration count; those centroids from
    a fixed-count run of the image); the stopping rule (shift after I - 1
    iterations > 1e-4 unless I = 1, reported shift = |c_I - c_(I-1)| and
    <= 1e-4 unless I = 100); segments equal the CPU union-find of the labels
    and its dense rank."""
    torch = torch_cuda
    ws = W.batch_workloads()
    say(f"full size batch_256: generating {len(ws)} images")
    planes = [torch.empty((3, w.n), dtype=torch.float32, device="cuda") for w in ws]
    sa.gen_images([w.image for w in ws], planes)
    imgs = []
    for w, P in zip(ws, planes):
        C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
        sa.uniform_random_initialize(P, C, w.init_seed)
        imgs.append(dict(points=P, width=w.image.width, height=w.image.height, centroids=C,
                         init=C.clone(),
                         labels=torch.full((w.n,), -7, dtype=torch.int32, device="cuda"),
                         segments=torch.full((w.n,), -7, dtype=torch.int32, device="cuda")))
    say(f"full size batch_256: one segment_batch call over "
        f"{sum(w.n for w in ws) / 1e6:.1f} Mpx")
    res = sa.segment_batch(imgs, max_iter=100, threshold=1e-4, connectivity=8)
    torch.cuda.synchronize()
    iters = np.array([r.iterations for r in res])
    assert len(set(iters.tolist())) > 1
    ctemp = torch.empty(sa.lib().split_hip_cluster_workspace(max(w.n for w in ws), 64, 3),
                        dtype=torch.uint8, device="cuda")
    for i, (w, im, r) in enumerate(zip(ws, imgs, res)):
        if i % 32 == 0:
            say(f"full size batch_256: checking images {i}..{min(i + 31, len(ws) - 1)}")
        wd, ht = im["width"], im["height"]
        tag = f"image {i} {wd}x{ht} k={w.k} I={r.iterations}"
        assert 1 <= r.iterations <= 100, tag
        # the centroids after I - 1 iterations, from a fixed-count run
        Cp = im["init"].clone()
        Lp = torch.empty(w.n, dtype=torch.int32, device="cuda")
        if r.iterations > 1:
            itp, shp = sa.cluster(im["points"], Cp, Lp, r.iterations - 1, -1.0, temp=ctemp,
                                  width=wd, height=ht)
            assert itp == r.iterations - 1 and shp > 1e-4, tag
        dense = torch.empty_like(Lp)
        sa.label_points(Cp, im["points"], dense, width=wd, height=ht)
        assert torch.equal(dense, im["labels"]), tag
        pts = im["points"].cpu().numpy()
        lab = im["labels"].cpu().numpy()
        cen = im["centroids"].cpu().numpy()
        means, cnt = _fp64_means(pts, lab, w.k)
        live = cnt > 0
        assert np.abs(cen[:, live].astype(np.float64) - means[:, live]).max() <= CENT_TOL, tag
        step = np.sqrt(((cen.astype(np.float64) - Cp.cpu().numpy()) ** 2).sum())
        assert abs(step - r.shift) <= 1e-6 * max(1.0, step), tag
        assert r.shift <= 1e-4 or r.iterations == 100, tag
        rep, ncomp = oracle.cc_uf(lab.reshape(ht, wd), 8)
        want, ns = oracle.compress(rep)
        assert r.num_segments == ns == ncomp, tag
        assert np.array_equal(im["segments"].cpu().numpy(), want), tag
