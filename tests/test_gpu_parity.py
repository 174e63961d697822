"""GPU parity of the gfx950 path against the CPU oracle, through the C ABI.

Bars (from the segmentation path's parity rules):
  * labels: identical to the float restatement of the reference's dense
    argmin (same arithmetic), and identical to the FP64 argmin except where the
    two nearest squared distances differ by < 1e-5 relative;
  * centroids: within 1e-5 (absolute, colours in [0, 1]) of the FP64 mean;
  * connected components: bit-exact (both sides emit the largest linear index
    of each component; the canonical min-index form is compared as well);
  * compress_labels: bit-exact dense rank.
"""
import numpy as np
import pytest

import oracle
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

CENT_TOL = 1e-5
TIE_REL = 1e-5


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def plain(torch_cuda):
    w = W.PLAIN_512
    img = W.make_image(w.image)
    return w, img


def seeded_centroids(img, k, seed):
    idx = oracle.sample_indices(seed, img.shape[1], k)
    return np.ascontiguousarray(img[:, idx])


# ----------------------------------------------------------- assignment ---

@pytest.mark.parametrize("k", [1, 3, 8, 32, 64, 65, 100, 256, 300, 1024])
def test_label_points_matches_reference_argmin(torch_cuda, plain, k):
    torch = torch_cuda
    w, img = plain
    say(f"label_points plain_512 k={k}")
    cent = seeded_centroids(img, k, 1000 + k)
    P, C = dev(torch, img), dev(torch, cent)
    L = torch.empty(img.shape[1], dtype=torch.int32, device="cuda")
    sa.label_points(C, P, L)
    got = host(L)
    # same float arithmetic as the reference's dense argmin: exact
    want = oracle.ref_label_points(cent, img)
    assert np.array_equal(got, want)
    # FP64 argmin up to near ties
    l64, gap = oracle.label_f64(img, cent)
    bad = (got != l64) & (gap >= TIE_REL)
    assert not bad.any(), f"{bad.sum()} labels differ beyond the near-tie rule"


@pytest.mark.parametrize("shape", [(1, 1), (1, 1000), (1000, 1), (17, 33), (37, 100),
                                   (64, 64), (129, 130), (512, 512)])
@pytest.mark.parametrize("k", [1, 8, 32, 65, 256, 1000])
def test_label_image_matches_reference_argmin(torch_cuda, shape, k):
    torch = torch_cuda
    h, w = shape
    say(f"label_image {w}x{h} k={k}")
    spec = W.ImageSpec(w, h, seed=77 + w + h, sites=max(1, w * h // 2000), shading=True,
                       noise_sigma=0.01)
    img = W.make_image(spec)
    rng = np.random.default_rng(k)
    cent = rng.random((3, k), dtype=np.float32)
    L = torch.empty(w * h, dtype=torch.int32, device="cuda")
    sa.label_points(dev(torch, cent), dev(torch, img), L, width=w, height=h)
    assert np.array_equal(host(L), oracle.ref_label_points(cent, img))


def test_cluster_image_plain_512(torch_cuda, plain):
    torch = torch_cuda
    w, img = plain
    say("cluster_image plain_512 k=8 x20")
    init = seeded_centroids(img, w.k, w.init_seed)
    C = dev(torch, init)
    L = torch.empty(img.shape[1], dtype=torch.int32, device="cuda")
    it, _ = sa.cluster(dev(torch, img), C, L, 20, -1.0, width=512, height=512)
    assert it == 20
    want_c, want_l = oracle.lloyd_f64(img, init, 20)
    got_c = host(C)
    assert np.abs(got_c - want_c).max() <= CENT_TOL
    c19, _ = oracle.lloyd_f64(img, init, 19)
    _, gap = oracle.label_f64(img, c19.astype(np.float32))
    assert not ((host(L) != want_l) & (gap >= TIE_REL)).any()


@pytest.mark.parametrize("shape,k,iters", [((512, 512), 32, 40), ((333, 517), 8, 30),
                                           ((256, 256), 256, 12), ((130, 97), 100, 15)])
def test_cluster_image_bounded_iterations(torch_cuda, shape, k, iters):
    """Iterations after the first skip pixels whose label is provably
    unchanged.  Their labels must still equal a dense assignment: the labels
    of an `iters` run equal label_image against the centroids of an
    `iters - 1` run (the trajectory is deterministic), and the centroids stay
    within 1e-5 of the FP64 trajectory."""
    torch = torch_cuda
    h, w = shape
    say(f"cluster_image bounded {w}x{h} k={k} x{iters}")
    spec = W.ImageSpec(w, h, seed=5 + w, sites=max(4, w * h // 3000), shading=True,
                       noise_sigma=0.01)
    img = W.make_image(spec)
    init = seeded_centroids(img, k, 900 + k)
    P = dev(torch, img)
    runs = {}
    for n_it in (iters - 1, iters):
        C = dev(torch, init)
        L = torch.empty(w * h, dtype=torch.int32, device="cuda")
        it, _ = sa.cluster(P, C, L, n_it, -1.0, width=w, height=h)
        assert it == n_it
        runs[n_it] = (host(C), host(L))
    c_prev, _ = runs[iters - 1]
    c_last, l_last = runs[iters]
    L2 = torch.empty(w * h, dtype=torch.int32, device="cuda")
    sa.label_points(dev(torch, c_prev), P, L2, width=w, height=h)
    assert np.array_equal(host(L2), l_last)
    assert np.array_equal(l_last, oracle.ref_label_points(c_prev, img))
    # teacher-forced: the centroids are the FP64 means of the GPU's own last
    # labels
    want_mean, _ = oracle.centroids_f64(l_last, img, k)
    assert np.abs(c_last - want_mean).max() <= CENT_TOL
    if k <= 32:
        # free-running FP64 trajectory (with many small clusters a single
        # near-tie flip moves a centroid by more than 1e-5, so only here)
        want_c, _ = oracle.lloyd_f64(img, init, iters)
        assert np.abs(c_last - want_c).max() <= CENT_TOL


@pytest.mark.parametrize("case", ["packed", "duplicates", "noise"])
def test_cluster_image_neighbour_walks(torch_cuda, case):
    """Bounded iterations at k > 64 re-assign failing pixels by walking the
    previous label's list of nearest centroids.  packed: 240 of 256 initial
    centroids sit within 0.004 of one colour, so pixels far from them exhaust
    the 32-entry lists and take the scan over every centroid; duplicates:
    pairs of equal centroids (ties to the lower index); noise: uniform
    colours, walks of every length.  The labels of an `iters` run must equal
    the dense argmin against the centroids of an `iters - 1` run."""
    torch = torch_cuda
    w, h, iters = 300, 211, 8
    rng = np.random.default_rng({"packed": 1, "duplicates": 2, "noise": 3}[case])
    if case == "noise":
        img = rng.random((3, w * h), dtype=np.float32)
        k = 256
        init = np.ascontiguousarray(img[:, rng.choice(w * h, k, replace=False)])
    else:
        spec = W.ImageSpec(w, h, seed=31, sites=40, shading=True, noise_sigma=0.02)
        img = W.make_image(spec)
        if case == "packed":
            k = 256
            init = np.empty((3, k), dtype=np.float32)
            init[:, :240] = 0.5 + 0.004 * rng.random((3, 240), dtype=np.float32)
            init[:, 240:] = img[:, rng.choice(w * h, 16, replace=False)]
        else:
            k = 96
            half = img[:, rng.choice(w * h, k // 2, replace=False)]
            init = np.ascontiguousarray(np.concatenate([half, half], axis=1))
    say(f"cluster_image neighbour walks {case} {w}x{h} k={k} x{iters}")
    P = dev(torch, img)
    runs = {}
    for n_it in (iters - 1, iters):
        C = dev(torch, init)
        L = torch.empty(w * h, dtype=torch.int32, device="cuda")
        it, _ = sa.cluster(P, C, L, n_it, -1.0, width=w, height=h)
        assert it == n_it
        runs[n_it] = (host(C), host(L))
    c_prev, _ = runs[iters - 1]
    c_last, l_last = runs[iters]
    assert np.array_equal(l_last, oracle.ref_label_points(c_prev, img))
    want_mean, _ = oracle.centroids_f64(l_last, img, k)
    assert np.abs(c_last - want_mean).max() <= CENT_TOL


def test_cluster_image_repeatable(torch_cuda, plain):
    torch = torch_cuda
    w, img = plain
    say("cluster_image repeatable")
    out = []
    for _ in range(2):
        C = dev(torch, seeded_centroids(img, 16, 4))
        L = torch.empty(img.shape[1], dtype=torch.int32, device="cuda")
        sa.cluster(dev(torch, img), C, L, 25, -1.0, width=512, height=512)
        out.append((host(C), host(L)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("n", [1, 5, 255, 256, 257, 1000, 4097])
def test_label_points_ragged_sizes(torch_cuda, n):
    torch = torch_cuda
    say(f"label_points ragged n={n}")
    rng = np.random.default_rng(n)
    img = rng.random((3, n), dtype=np.float32)
    cent = rng.random((3, 7), dtype=np.float32)
    L = torch.empty(n, dtype=torch.int32, device="cuda")
    sa.label_points(dev(torch, cent), dev(torch, img), L)
    assert np.array_equal(host(L), oracle.ref_label_points(cent, img))


def test_label_points_unaligned_planes(torch_cuda):
    # n % 4 != 0 puts planes 1 and 2 off 16-byte alignment: scalar path
    torch = torch_cuda
    say("label_points unaligned n=262147")
    rng = np.random.default_rng(7)
    n = 262147
    img = rng.random((3, n), dtype=np.float32)
    cent = rng.random((3, 16), dtype=np.float32)
    L = torch.empty(n, dtype=torch.int32, device="cuda")
    sa.label_points(dev(torch, cent), dev(torch, img), L)
    assert np.array_equal(host(L), oracle.ref_label_points(cent, img))


def test_label_points_ties_and_duplicates(torch_cuda):
    torch = torch_cuda
    say("label_points ties")
    # duplicate centroids: the lower index takes every pixel
    img = np.random.default_rng(3).random((3, 4096), dtype=np.float32)
    c = np.random.default_rng(4).random((3, 5), dtype=np.float32)
    c[:, 3] = c[:, 1]
    L = torch.empty(4096, dtype=torch.int32, device="cuda")
    sa.label_points(dev(torch, c), dev(torch, img), L)
    got = host(L)
    assert not (got == 3).any()
    assert np.array_equal(got, oracle.ref_label_points(c, img))
    # pixels equal to a centroid get that centroid
    c2 = np.ascontiguousarray(img[:, [10, 20, 30]])
    sa.label_points(dev(torch, c2), dev(torch, img), L)
    got = host(L)
    assert got[10] == 0 and got[20] == 1 and got[30] == 2


# -------------------------------------------------------------- centroids ---

@pytest.mark.parametrize("k", [1, 8, 32, 256, 1024])
def test_calculate_centroids_fp64(torch_cuda, plain, k):
    torch = torch_cuda
    w, img = plain
    say(f"calculate_centroids plain_512 k={k}")
    cent = seeded_centroids(img, min(k, 1024), 2000 + k)
    labels = oracle.ref_label_points(cent, img)
    out = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.calculate_centroids(dev(torch, labels), dev(torch, img), out)
    got = host(out)
    want, counts = oracle.centroids_f64(labels, img, k)
    assert np.abs(got - want).max() <= CENT_TOL
    assert np.all(got[:, counts == 0] == 0)


def test_calculate_centroids_empty_and_out_of_range(torch_cuda):
    torch = torch_cuda
    say("calculate_centroids empty/out-of-range labels")
    rng = np.random.default_rng(11)
    n, k = 10000, 6
    img = rng.random((3, n), dtype=np.float32)
    labels = rng.integers(-2, 9, n).astype(np.int32)
    labels[labels == 2] = 3  # label 2 empty
    out = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.calculate_centroids(dev(torch, labels), dev(torch, img), out)
    got = host(out)
    want, counts = oracle.centroids_f64(labels, img, k)
    assert counts[2] == 0
    assert np.all(got[:, 2] == 0)
    assert np.abs(got - want).max() <= CENT_TOL


def test_calculate_centroids_closed_forms(torch_cuda):
    torch = torch_cuda
    say("calculate_centroids closed forms")
    n, k = 3 * 4096 + 5, 3
    # constant image -> constant centroid
    img = np.full((3, n), 0.25, dtype=np.float32)
    labels = (np.arange(n) % k).astype(np.int32)
    out = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.calculate_centroids(dev(torch, labels), dev(torch, img), out)
    assert np.all(host(out) == 0.25)
    # affine ramp x_i = i/n with labels i mod k: mean of {j, j+k, ...}/n
    ramp = (np.arange(n, dtype=np.float64) / n).astype(np.float32)
    img = np.stack([ramp, 1 - ramp, ramp * 0.5]).astype(np.float32)
    sa.calculate_centroids(dev(torch, labels), dev(torch, img), out)
    got = host(out)
    for j in range(k):
        idx = np.arange(j, n, k)
        assert abs(got[0, j] - img[0, idx].astype(np.float64).mean()) <= CENT_TOL


# ------------------------------------------------------------------ Lloyd ---

def test_cluster_fixed_iterations_plain_512(torch_cuda, plain):
    torch = torch_cuda
    w, img = plain
    say("cluster plain_512 k=8 x20")
    init = seeded_centroids(img, w.k, w.init_seed)
    C = dev(torch, init)
    L = torch.empty(img.shape[1], dtype=torch.int32, device="cuda")
    it, shift = sa.cluster(dev(torch, img), C, L, w.max_iter, w.threshold)
    assert it == 20
    got_c, got_l = host(C), host(L)
    want_c, want_l = oracle.lloyd_f64(img, init, 20)
    assert np.abs(got_c - want_c).max() <= CENT_TOL
    # labels of the last assignment: FP64 labels against the FP64 centroids
    # after 19 updates
    c19, _ = oracle.lloyd_f64(img, init, 19)
    l64, gap = oracle.label_f64(img, c19.astype(np.float32))
    bad = (got_l != want_l) & (gap >= TIE_REL)
    assert not bad.any()
    # teacher-forced: one GPU step from the GPU's own centroids equals the
    # reference arithmetic exactly for labels and FP64 within 1e-5
    L2 = torch.empty_like(L)
    sa.label_points(C, dev(torch, img), L2)
    lab = host(L2)
    assert np.array_equal(lab, oracle.ref_label_points(got_c, img))
    C2 = torch.empty_like(C)
    sa.calculate_centroids(L2, dev(torch, img), C2)
    assert np.abs(host(C2) - oracle.centroids_f64(lab, img, w.k)[0]).max() <= CENT_TOL


@pytest.mark.parametrize("tiled", [False, True])
def test_cluster_stops_on_threshold(torch_cuda, plain, tiled):
    torch = torch_cuda
    w, img = plain
    say(f"cluster plain_512 to tolerance tiled={tiled}")
    init = seeded_centroids(img, 8, 99)
    C = dev(torch, init)
    L = torch.empty(img.shape[1], dtype=torch.int32, device="cuda")
    geo = dict(width=512, height=512) if tiled else {}
    it, shift = sa.cluster(dev(torch, img), C, L, 100, 1e-4, **geo)
    assert it > 2
    assert 1 <= it <= 100
    assert shift <= 1e-4 or it == 100
    # the FP64 trajectory reaches the same centroids after the same count
    want_c, _ = oracle.lloyd_f64(img, init, it)
    assert np.abs(host(C) - want_c).max() <= CENT_TOL


def test_cluster_zero_iterations_leaves_inputs(torch_cuda, plain):
    torch = torch_cuda
    w, img = plain
    say("cluster max_iter=0")
    init = seeded_centroids(img, 8, 5)
    C = dev(torch, init)
    L = torch.full((img.shape[1],), -7, dtype=torch.int32, device="cuda")
    it, _ = sa.cluster(dev(torch, img), C, L, 0, -1.0)
    assert it == 0
    assert np.array_equal(host(C), init)
    assert np.all(host(L) == -7)


def test_uniform_random_initialize_matches_sampler(torch_cuda, plain):
    torch = torch_cuda
    w, img = plain
    say("uniform_random_initialize")
    for k, seed in [(8, 1), (32, 2), (256, 3), (1024, 4)]:
        C = torch.empty((3, k), dtype=torch.float32, device="cuda")
        sa.uniform_random_initialize(dev(torch, img), C, seed)
        idx = oracle.sample_indices(seed, img.shape[1], k)
        assert len(set(idx.tolist())) == k
        assert np.array_equal(host(C), img[:, idx])


# ---------------------------------------------------- connected components ---

def run_cc(torch, lab_hw, conn):
    h, w = lab_hw.shape
    out = torch.empty(h * w, dtype=torch.int32, device="cuda")
    sa.connected_components(dev(torch, lab_hw), out, connectivity=conn)
    return host(out)


@pytest.mark.parametrize("conn", [4, 8])
@pytest.mark.parametrize("shape", [(1, 1), (1, 77), (77, 1), (31, 33), (64, 64),
                                   (100, 37), (257, 129), (63, 65), (65, 129),
                                   (1, 130), (130, 1), (200, 300)])
def test_cc_random_labels(torch_cuda, conn, shape):
    torch = torch_cuda
    say(f"cc random shape={shape} conn={conn}")
    rng = np.random.default_rng(hash(shape) % 1000 + conn)
    for nlab in (2, 3, 5):
        lab = rng.integers(0, nlab, shape).astype(np.int32)
        got = run_cc(torch, lab, conn)
        want, _ = oracle.cc_uf(lab, conn)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("conn", [4, 8])
def test_cc_voronoi_labels_and_reference_fixed_point(torch_cuda, plain, conn):
    torch = torch_cuda
    w, img = plain
    say(f"cc plain_512 cluster labels conn={conn}")
    cent = seeded_centroids(img, 8, 7)
    lab = oracle.ref_label_points(cent, img).reshape(512, 512)
    got = run_cc(torch, lab, conn)
    want, _ = oracle.cc_uf(lab, conn)
    assert np.array_equal(got, want)
    assert np.array_equal(oracle.canonical_min(got), oracle.canonical_min(want))
    if conn == 4:
        ref, it = oracle.ref_connected_components(lab, 100000)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("conn", [4, 8])
def test_cc_closed_forms(torch_cuda, conn):
    torch = torch_cuda
    say(f"cc closed forms conn={conn}")
    h, w = 96, 130
    uni = np.zeros((h, w), np.int32)
    assert np.all(run_cc(torch, uni, conn) == h * w - 1)
    cb = ((np.arange(h)[:, None] + np.arange(w)[None, :]) & 1).astype(np.int32)
    got = run_cc(torch, cb, conn)
    n_comp = len(np.unique(got))
    assert n_comp == (h * w if conn == 4 else 2)
    stripes = np.repeat((np.arange(h) % 2)[:, None], w, axis=1).astype(np.int32)
    assert len(np.unique(run_cc(torch, stripes, conn))) == (h if conn == 4 else h)


@pytest.mark.parametrize("conn", [4, 8])
def test_cc_spiral_stress_2048(torch_cuda, conn):
    torch = torch_cuda
    say(f"cc spiral/checkerboard 2048 conn={conn}")
    lab = W.make_cc_stress(2048, 2048)
    got = run_cc(torch, lab, conn)
    want, ncomp = oracle.cc_uf(lab, conn)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (f"{bad.size} pixels differ; first {bad[:5].tolist()} "
                           f"got {got[bad[:5]].tolist()} want {want[bad[:5]].tolist()}")


def test_cc_spiral_repeated_runs_agree(torch_cuda):
    """Regression: racing stores on one forest word (a root-jump walk's
    path halving against the word's own final store) must not leave a tile
    root short of its root; every run of the long-chain spiral image is
    exact, with and without compression."""
    torch = torch_cuda
    lab = W.make_cc_stress(2048, 2048)
    L = dev(torch, lab)
    out = torch.empty(lab.size, dtype=torch.int32, device="cuda")
    for conn in (4, 8):
        want, _ = oracle.cc_uf(lab, conn)
        dense, cnt = oracle.compress(want)
        for rep in range(4):
            say(f"cc spiral 2048 conn={conn} run {rep}")
            sa.connected_components(L, out, connectivity=conn)
            got = host(out)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, (conn, rep, bad.size, bad[:5].tolist())
            assert sa.compress_labels(out) == cnt
            assert np.array_equal(host(out), dense)


# ------------------------------------------------------------- compress ---

def test_compress_labels_matches_sort_unique(torch_cuda):
    torch = torch_cuda
    say("compress_labels random")
    rng = np.random.default_rng(5)
    for lo, hi, n in [(0, 10, 1000), (-50, 50, 7777), (-1000000, 1000000, 100000),
                      (5, 6, 1), (2**31 - 100, 2**31 - 1, 3000)]:
        v = rng.integers(lo, hi, n, dtype=np.int64).astype(np.int32)
        t = dev(torch, v)
        # a value span wider than 32 N needs a temp sized for it
        span = int(v.max()) - int(v.min()) + 1
        tmp = torch.empty(max(sa.lib().split_hip_compress_labels_workspace(n, span), 4 * n),
                          dtype=torch.uint8, device="cuda")
        cnt = sa.compress_labels(t, temp=tmp)
        want, wcnt = oracle.compress(v)
        assert cnt == wcnt
        assert np.array_equal(host(t), want)
        assert np.array_equal(want, np.unique(v, return_inverse=True)[1].ravel())


def test_compress_labels_range_error(torch_cuda):
    torch = torch_cuda
    say("compress_labels range error")
    v = np.array([-2**31, 2**31 - 1, 0, 5], dtype=np.int32)
    t = dev(torch, v)
    tmp = torch.empty(64 + 4 * 4 + 64, dtype=torch.uint8, device="cuda")
    with pytest.raises(sa.SplitError) as e:
        sa.compress_labels(t, temp=tmp)
    assert e.value.status == 4
    assert np.array_equal(host(t), v)


# ------------------------------------------------------------- pipeline ---

def test_segment_plain_512(torch_cuda, plain):
    torch = torch_cuda
    w, img = plain
    say("segment plain_512")
    init = seeded_centroids(img, w.k, w.init_seed)
    C = dev(torch, init)
    L = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    r = sa.segment(dev(torch, img), 512, 512, C, L, S, max_iter=20, threshold=-1.0,
                   connectivity=4)
    lab = host(L).reshape(512, 512)
    rep, ncomp = oracle.cc_uf(lab, 4)
    want, cnt = oracle.compress(rep)
    assert r.iterations == 20
    assert r.num_segments == cnt == ncomp
    assert np.array_equal(host(S), want)


@pytest.mark.parametrize("n,k,iters", [(262144 - 7, 8, 20), (4097, 5, 15), (17, 3, 6),
                                       (100_003, 32, 25)])
def test_cluster_points_without_width(torch_cuda, n, k, iters):
    """cluster() of a point set with no image width: the bounded tiled path
    walks it as rows of 16 points, the last row partial when n is not a
    multiple of 16.  Centroids within 1e-5 of the FP64 trajectory; labels
    equal the FP64 argmin against the previous centroids except near ties,
    and equal a dense assignment against the (iters - 1)-run's centroids."""
    torch = torch_cuda
    say(f"cluster points n={n} k={k} x{iters} (no width)")
    spec = W.ImageSpec(512, 512, seed=0x9A66 + n, sites=64, noise_sigma=0.02)
    img = np.ascontiguousarray(W.make_image(spec)[:, :n])
    init = seeded_centroids(img, k, 0x1A66 + k)
    P = dev(torch, img)
    runs = {}
    for n_it in (iters - 1, iters):
        C = dev(torch, init)
        L = torch.empty(n, dtype=torch.int32, device="cuda")
        it, _ = sa.cluster(P, C, L, n_it, -1.0)
        assert it == n_it
        runs[n_it] = (host(C), host(L))
    c_prev, _ = runs[iters - 1]
    c_last, l_last = runs[iters]
    want_c, want_l = oracle.lloyd_f64(img, init, iters)
    assert np.abs(c_last - want_c).max() <= CENT_TOL
    cp, _ = oracle.lloyd_f64(img, init, iters - 1)
    _, gap = oracle.label_f64(img, cp.astype(np.float32))
    assert not ((l_last != want_l) & (gap >= TIE_REL)).any()
    L2 = torch.empty(n, dtype=torch.int32, device="cuda")
    sa.label_points(dev(torch, c_prev), P, L2)
    assert np.array_equal(host(L2), l_last)


@pytest.mark.parametrize("k", [8, 300])
def test_cluster_runs_path(torch_cuda, plain, k):
    """The 1-D runs path (float sums): k > 256, or a temp block too small for
    the tiled path's per-pixel state.  Centroids within 1e-5 of the FP64
    means of the returned labels; labels equal a dense assignment against
    the previous iteration's centroids."""
    torch = torch_cuda
    w, img = plain
    n = img.shape[1]
    iters = 8
    say(f"cluster runs path plain_512 k={k} x{iters}")
    init = seeded_centroids(img, k, 0x2B + k)
    P = dev(torch, img)
    # k = 8: a temp that holds the runs path's partial columns but not the
    # tiled path's 20 bytes per pixel
    temp = torch.empty(1 << 20, dtype=torch.uint8, device="cuda") if k == 8 else None
    runs = {}
    for n_it in (iters - 1, iters):
        C = dev(torch, init)
        L = torch.empty(n, dtype=torch.int32, device="cuda")
        it, _ = sa.cluster(P, C, L, n_it, -1.0, temp=temp)
        assert it == n_it
        runs[n_it] = (host(C), host(L))
    L2 = torch.empty(n, dtype=torch.int32, device="cuda")
    sa.label_points(dev(torch, runs[iters - 1][0]), P, L2)
    assert np.array_equal(host(L2), runs[iters][1])
    m, cnt = oracle.centroids_f64(runs[iters][1], img, k)
    live = cnt > 0
    assert np.abs(runs[iters][0][:, live] - m[:, live]).max() <= CENT_TOL
