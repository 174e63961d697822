"""CPU pins of the oracle (no GPU).

The reference has no tests, fixtures or golden vectors for this path, so each
oracle function is pinned by closed forms, brute force on tiny inputs or an
independent implementation, chosen so that a dropped term, a wrong index or a
transposed operand fails:

  oracle_label_f64          brute-force numpy argmin (float64), K=1, equal
                            pixels, duplicate centroids, bisector ties
  ref_label_points          brute-force numpy float32 argmin with the same
                            fused multiply-add rounding; equals the FP64 argmin
                            up to near ties
  oracle_centroids_f64      numpy bincount means; constant image; ramp with
                            labels i mod K (analytic); single pixel per label
  ref_calculate_centroids   equals FP64 means when no label is empty; the
                            empty-label slot shift of the reference
                            (centroids.cu:78-100) on a hand-worked case
  oracle_lloyd_f64          inertia non-increasing; fixed point idempotent;
                            one step == label + mean
  ref_cluster               threshold < 0 runs exactly max_iter; with no
                            aliasing (k <= 2, tiny n) the shift rule reduces to
                            |delta - old_delta|
  oracle_cc_uf              scipy.ndimage.label and a pure-Python BFS;
                            uniform / checkerboard / stripes / spiral closed
                            forms
  ref_connected_components  equals union-find at convergence; iteration count
                            of a 1-px spiral grows with its turns; truncation
                            by max_iter
  oracle_compress           numpy.unique(return_inverse)
  oracle_sample_indices     SplitMix64 reference values (published test vector
                            of the generator) and distinctness
"""
import collections

import numpy as np
import pytest
from scipy import ndimage

import oracle
from split_amd import workloads as W


def fma32(a, b, c):
    """float32 fused multiply-add via exact float64 product and a single
    rounding (exact: the float32 product fits in float64's 53 bits, and the
    sum is rounded once to float32 after an exact-enough double add — a
    double-rounding case needs a tie at 2^-24 precision, absent here)."""
    return (np.float64(a) * np.float64(b) + np.float64(c)).astype(np.float32)


def brute_ref_labels(cent, pts):
    k = cent.shape[1]
    d = np.empty((k, pts.shape[1]), np.float32)
    for j in range(k):
        dx = (pts[0] - cent[0, j]).astype(np.float32)
        dy = (pts[1] - cent[1, j]).astype(np.float32)
        dz = (pts[2] - cent[2, j]).astype(np.float32)
        d[j] = fma32(dz, dz, fma32(dy, dy, (dx * dx).astype(np.float32)))
    return np.argmin(d, axis=0).astype(np.int32)  # argmin: first minimum


@pytest.fixture(scope="module")
def rng():
    return np.random.default_rng(12345)


# ------------------------------------------------------------- labelling ---

def test_label_f64_matches_bruteforce(rng):
    pts = rng.random((3, 500), dtype=np.float32)
    cent = rng.random((3, 7), dtype=np.float32)
    lab, gap = oracle.label_f64(pts, cent)
    d = ((pts.astype(np.float64)[:, None, :] - cent.astype(np.float64)[:, :, None]) ** 2).sum(0)
    assert np.array_equal(lab, d.argmin(0))
    s = np.sort(d, axis=0)
    assert np.allclose(gap, (s[1] - s[0]) / s[1])


def test_label_closed_forms(rng):
    pts = rng.random((3, 64), dtype=np.float32)
    # K = 1: every label is 0, gap 1
    lab, gap = oracle.label_f64(pts, pts[:, :1])
    assert np.all(lab == 0) and np.all(gap == 1)
    # pixels equal to centroids get them (gap > 0 unless duplicated)
    cent = np.ascontiguousarray(pts[:, [3, 9, 17]])
    lab, _ = oracle.label_f64(pts, cent)
    assert lab[3] == 0 and lab[9] == 1 and lab[17] == 2
    # duplicate centroids: the lower index wins, gap 0 for its pixels
    cent = np.ascontiguousarray(np.stack([cent[:, 0], cent[:, 1], cent[:, 1]], 1))
    lab, gap = oracle.label_f64(pts, cent)
    assert not (lab == 2).any()
    assert np.all(gap[lab == 1] == 0)
    # pixels on the bisector of two centroids: the lower index
    c2 = np.array([[0.25, 0.75], [0.5, 0.5], [0.5, 0.5]], np.float32)
    p2 = np.array([[0.5], [0.1], [0.9]], np.float32)
    assert oracle.label_f64(p2, c2)[0][0] == 0
    assert oracle.ref_label_points(c2, p2)[0] == 0


def test_ref_label_points_bruteforce_fma(rng):
    pts = rng.random((3, 2000), dtype=np.float32)
    cent = rng.random((3, 13), dtype=np.float32)
    assert np.array_equal(oracle.ref_label_points(cent, pts), brute_ref_labels(cent, pts))


def test_ref_label_points_vs_fp64_near_ties(rng):
    pts = rng.random((3, 20000), dtype=np.float32)
    cent = rng.random((3, 32), dtype=np.float32)
    r = oracle.ref_label_points(cent, pts)
    l64, gap = oracle.label_f64(pts, cent)
    assert not ((r != l64) & (gap >= 1e-5)).any()


# ------------------------------------------------------------- centroids ---

def test_centroids_f64_bincount(rng):
    n, k = 5000, 9
    pts = rng.random((3, n), dtype=np.float32)
    lab = rng.integers(-1, k + 2, n).astype(np.int32)
    got, counts = oracle.centroids_f64(lab, pts, k)
    ok = (lab >= 0) & (lab < k)
    cnt = np.bincount(lab[ok], minlength=k)
    assert np.array_equal(counts, cnt)
    for c in range(3):
        s = np.bincount(lab[ok], weights=pts[c, ok].astype(np.float64), minlength=k)
        want = np.where(cnt > 0, s / np.maximum(cnt, 1), 0.0)
        assert np.allclose(got[c], want, rtol=0, atol=1e-12)


def test_centroids_closed_forms():
    n, k = 3001, 3
    const = np.full((3, n), 0.375, np.float32)
    lab = (np.arange(n) % k).astype(np.int32)
    assert np.all(oracle.centroids_f64(lab, const, k)[0] == 0.375)
    # x_i = i: mean of {j, j+k, ...} below n is j + k*(m-1)/2
    ramp = np.stack([np.arange(n), 2 * np.arange(n), np.zeros(n)]).astype(np.float32)
    got, counts = oracle.centroids_f64(lab, ramp, k)
    for j in range(k):
        m = len(range(j, n, k))
        assert got[0, j] == j + k * (m - 1) / 2
        assert got[1, j] == 2 * (j + k * (m - 1) / 2)
    # one pixel per label: the pixel itself
    pts = np.arange(12, dtype=np.float32).reshape(3, 4) / 16
    got, _ = oracle.centroids_f64(np.arange(4, dtype=np.int32), pts, 4)
    assert np.array_equal(got, pts)


def test_ref_calculate_centroids_equals_mean_without_empty(rng):
    n, k = 4000, 6
    pts = rng.random((3, n), dtype=np.float32)
    lab = rng.integers(0, k, n).astype(np.int32)
    ref = oracle.ref_calculate_centroids(lab, pts, k)
    want, _ = oracle.centroids_f64(lab, pts, k)
    assert np.abs(ref - want).max() < 1e-5


def test_ref_calculate_centroids_empty_slot_shift():
    # labels {0, 2} with k = 3: reduce_by_key writes the sums of label 2 into
    # slot 1 (compacted); slot 1 is multiplied by 1/count_1 = 1/0 = inf and
    # slot 2 (never written, 0) by 1/count_2.
    pts = np.array([[1.0, 3.0, 5.0]] * 3, np.float32)
    lab = np.array([0, 0, 2], np.int32)
    ref = oracle.ref_calculate_centroids(lab, pts, 3)
    assert ref[0, 0] == 2.0
    assert np.isinf(ref[0, 1])
    assert ref[0, 2] == 0.0


# ----------------------------------------------------------------- Lloyd ---

def test_lloyd_f64_inertia_monotone_and_fixed_point(rng):
    w = W.scaled(W.PLAIN_512, 128, 96)
    img = W.make_image(w.image)
    init = np.ascontiguousarray(img[:, oracle.sample_indices(3, w.n, 6)])
    prev = np.inf
    for it in range(1, 12):
        c, lab = oracle.lloyd_f64(img, init, it)
        # inertia of the last assignment against the updated centroids
        e = oracle.inertia(img, c.astype(np.float32), lab)
        assert e <= prev * (1 + 1e-6)
        prev = e
    # at a fixed point another iteration changes nothing (same trajectory)
    c200, l200 = oracle.lloyd_f64(img, init, 200)
    c201, l201 = oracle.lloyd_f64(img, init, 201)
    assert np.array_equal(l200, l201)
    assert np.array_equal(c200, c201)


def test_lloyd_one_step_is_label_then_mean(rng):
    pts = rng.random((3, 3000), dtype=np.float32)
    init = np.ascontiguousarray(pts[:, :5])
    c, lab = oracle.lloyd_f64(pts, init, 1)
    l0, _ = oracle.label_f64(pts, init)
    assert np.array_equal(lab, l0)
    assert np.array_equal(c, oracle.centroids_f64(l0, pts, 5)[0])


def test_ref_cluster_fixed_iterations(rng):
    pts = rng.random((3, 3000), dtype=np.float32)
    init = np.ascontiguousarray(pts[:, :4])
    cent, lab, it, _ = oracle.ref_cluster(pts, init, 7, -1.0)
    assert it == 7
    c = init.copy()
    for _ in range(7):
        l = oracle.ref_label_points(c, pts)
        c = oracle.ref_calculate_centroids(l, pts, 4)
    assert np.array_equal(cent, c) and np.array_equal(lab, l)


def test_ref_cluster_stopping_rule_without_aliasing():
    # n = 1, k = 1: no distance entry reaches the saved copy (2N + j >= N*K),
    # but the reciprocal count 1/1 overwrites its first value.  Iteration 1:
    # saved (0, 0, 0) -> (1, 0, 0), new (0.25, 0.5, 0.75), delta^2 = 0.75^2 +
    # 0.5^2 + 0.75^2.  Iteration 2: saved (0.25, 0.5, 0.75) -> (1, 0.5, 0.75),
    # delta = 0.75.  Iteration 3 repeats 0.75, |delta - old_delta| = 0: stop.
    pts = np.array([[0.25], [0.5], [0.75]], np.float32)
    init = np.zeros((3, 1), np.float32)
    _, _, it, delta = oracle.ref_cluster(pts, init, 1, 1e-3)
    assert it == 1 and np.isclose(delta, np.sqrt(0.75**2 + 0.5**2 + 0.75**2))
    _, _, it, delta = oracle.ref_cluster(pts, init, 10, 1e-3)
    assert it == 3 and np.isclose(delta, 0.75)


# ------------------------------------------------------------------- CC ---

def bfs_components(lab, conn):
    h, w = lab.shape
    seen = np.full((h, w), -1, np.int64)
    out = np.empty(h * w, np.int64)
    nb4 = [(-1, 0), (1, 0), (0, -1), (0, 1)]
    nb = nb4 + ([(-1, -1), (-1, 1), (1, -1), (1, 1)] if conn == 8 else [])
    comp = 0
    for y in range(h):
        for x in range(w):
            if seen[y, x] >= 0:
                continue
            q = collections.deque([(y, x)])
            seen[y, x] = comp
            members = []
            while q:
                cy, cx = q.popleft()
                members.append(cy * w + cx)
                for dy, dx in nb:
                    ny, nx = cy + dy, cx + dx
                    if 0 <= ny < h and 0 <= nx < w and seen[ny, nx] < 0 \
                            and lab[ny, nx] == lab[cy, cx]:
                        seen[ny, nx] = comp
                        q.append((ny, nx))
            out[members] = max(members)
            comp += 1
    return out, comp


@pytest.mark.parametrize("conn", [4, 8])
def test_cc_uf_bfs(rng, conn):
    for shape in [(1, 1), (1, 9), (9, 1), (7, 11), (20, 13)]:
        for nl in (2, 3):
            lab = rng.integers(0, nl, shape).astype(np.int32)
            got, cnt = oracle.cc_uf(lab, conn)
            want, wcnt = bfs_components(lab, conn)
            assert np.array_equal(got, want) and cnt == wcnt


@pytest.mark.parametrize("conn", [4, 8])
def test_cc_uf_scipy(rng, conn):
    lab = rng.integers(0, 3, (200, 170)).astype(np.int32)
    got, cnt = oracle.cc_uf(lab, conn)
    st = ndimage.generate_binary_structure(2, 1 if conn == 4 else 2)
    total = 0
    canon = oracle.canonical_min(got)
    for v in np.unique(lab):
        comp, m = ndimage.label(lab == v, structure=st)
        total += m
        # every scipy component is one of ours
        mask = (lab == v).ravel()
        ids = comp.ravel()[mask]
        ours = canon[mask]
        pairs = set(zip(ids.tolist(), ours.tolist()))
        assert len(pairs) == m
    assert cnt == total


@pytest.mark.parametrize("conn", [4, 8])
def test_cc_closed_forms(conn):
    h, w = 30, 41
    assert oracle.cc_uf(np.zeros((h, w), np.int32), conn)[1] == 1
    cb = ((np.arange(h)[:, None] + np.arange(w)[None, :]) & 1).astype(np.int32)
    assert oracle.cc_uf(cb, conn)[1] == (h * w if conn == 4 else 2)
    stripes = np.repeat((np.arange(h) % 2)[:, None], w, 1).astype(np.int32)
    assert oracle.cc_uf(stripes, conn)[1] == h


def test_spiral_stress_components():
    # one square spiral (wall + corridor) on top, checkerboard band below:
    # 8-connectivity joins the band's two colours with the spiral parts
    lab = W.make_cc_stress(64, 64)
    top = lab[:56, :56]
    assert oracle.cc_uf(np.ascontiguousarray(top), 4)[1] == 2
    assert oracle.cc_uf(np.ascontiguousarray(top), 8)[1] == 2
    band = np.ascontiguousarray(lab[56:])
    assert oracle.cc_uf(band, 4)[1] == band.size
    assert oracle.cc_uf(band, 8)[1] == 2


def test_ref_cc_equals_uf_at_convergence(rng):
    for shape in [(1, 1), (5, 8), (33, 17), (64, 64)]:
        lab = rng.integers(0, 3, shape).astype(np.int32)
        ref, it = oracle.ref_connected_components(lab, 100000)
        uf, _ = oracle.cc_uf(lab, 4)
        assert np.array_equal(ref, uf)


def test_ref_cc_iterations_grow_with_spiral_and_truncate():
    small = np.ascontiguousarray(W.make_cc_stress(64, 64)[:56, :56])
    big = np.ascontiguousarray(W.make_cc_stress(256, 256)[:224, :224])
    _, it_s = oracle.ref_connected_components(small, 100000)
    _, it_b = oracle.ref_connected_components(big, 100000)
    assert it_b > 2 * it_s
    trunc, it = oracle.ref_connected_components(big, 10)
    assert it == 10
    assert not np.array_equal(trunc, oracle.cc_uf(big, 4)[0])


# ------------------------------------------------------------- compress ---

def test_compress_numpy_unique(rng):
    for lo, hi, n in [(0, 5, 100), (-1000, 1000, 5000), (-2**31, 2**31 - 1, 3000)]:
        v = rng.integers(lo, hi, n, dtype=np.int64).astype(np.int32)
        got, cnt = oracle.compress(v)
        u, inv = np.unique(v, return_inverse=True)
        assert cnt == len(u) and np.array_equal(got, inv.ravel())


# -------------------------------------------------------------- sampler ---

def test_splitmix64_sequence():
    # SplitMix64 from state 0: first outputs 0xE220A8397B1DCDAF,
    # 0x6E789E6AA1B965F4 (the generator's published reference values).
    # index = high 64 bits of draw * n; with n = 2^32 that is draw >> 32.
    idx = oracle.sample_indices(0, 2**32, 2)
    assert idx[0] == 0xE220A8397B1DCDAF >> 32
    assert idx[1] == 0x6E789E6AA1B965F4 >> 32


def test_sampler_distinct_and_redraws():
    idx = oracle.sample_indices(7, 10, 10)
    assert sorted(idx.tolist()) == list(range(10))
    idx = oracle.sample_indices(7, 1000, 100)
    assert len(set(idx.tolist())) == 100 and idx.min() >= 0 and idx.max() < 1000
