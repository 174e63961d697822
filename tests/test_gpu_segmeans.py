"""Segment means (calculate_centroids with k = number of segments, beyond the
register-resident kernels) and propagate_centroids, through the C ABI.

Bars: centroids within 1e-5 of the FP64 per-label mean (the quantised
integer sums move a mean by at most max|v| * 2^-22); a label with no pixel
gets the zero centroid; labels outside [0, k) are ignored; propagation is an
exact gather.
"""
import numpy as np
import pytest

import oracle
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _means_f64(labels, img, k):
    lab = labels.astype(np.int64)
    ok = (lab >= 0) & (lab < k)
    cnt = np.bincount(lab[ok], minlength=k)
    out = np.zeros((3, k))
    for c in range(3):
        s = np.bincount(lab[ok], weights=img[c][ok].astype(np.float64), minlength=k)
        out[c] = np.where(cnt > 0, s / np.maximum(cnt, 1), 0.0)
    return out, cnt


@pytest.mark.parametrize("n,k", [(1_000_000, 100_000), (999_999, 5_000), (4097, 2_000)])
def test_segment_means_random_labels(torch_cuda, n, k):
    torch = torch_cuda
    say(f"segment means: random labels n={n} k={k}")
    rng = np.random.default_rng(n + k)
    img = rng.random((3, n), dtype=np.float32)
    # runs of equal labels (as connected components give), some out of range
    runs = rng.integers(1, 40, size=n)
    lab = np.repeat(rng.integers(-2, k + 3, size=n), runs)[:n].astype(np.int32)
    lab[(lab == k // 2) | (lab == k - 1)] = -1  # two labels with no pixel
    C = torch.full((3, k), -5.0, dtype=torch.float32, device="cuda")
    sa.calculate_centroids(_dev(torch, lab), _dev(torch, img), C)
    got = C.cpu().numpy()
    want, cnt = _means_f64(lab, img, k)
    assert np.abs(got - want).max() <= TOL
    assert (cnt == 0).any()
    assert np.all(got[:, cnt == 0] == 0.0)


def test_segment_means_of_components(torch_cuda):
    torch = torch_cuda
    w = W.SHADED_4096
    say("segment means of shaded_4096 components")
    img = W.make_image(w.image)
    P = _dev(torch, img)
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    L = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    sa.uniform_random_initialize(P, C, w.init_seed)
    r = sa.segment(P, 4096, 4096, C, L, S, max_iter=10, threshold=-1.0, connectivity=8)
    k = r.num_segments
    assert k > 1024
    M = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.calculate_centroids(S, P, M)
    seg = S.cpu().numpy()
    want, cnt = _means_f64(seg, img, k)
    assert np.all(cnt > 0)
    assert np.abs(M.cpu().numpy() - want).max() <= TOL
    # propagate: every pixel takes its segment's mean
    out = P.clone()
    sa.propagate_centroids(S, M, out)
    m = M.cpu().numpy()
    assert np.array_equal(out.cpu().numpy(), m[:, seg])


def test_propagate_out_of_range_untouched(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(3)
    n, k = 10_001, 37
    pts = rng.random((3, n), dtype=np.float32)
    cent = rng.random((3, k), dtype=np.float32)
    lab = rng.integers(-3, k + 3, size=n).astype(np.int32)
    out = _dev(torch, pts)
    sa.propagate_centroids(_dev(torch, lab), _dev(torch, cent), out)
    got = out.cpu().numpy()
    ok = (lab >= 0) & (lab < k)
    assert np.array_equal(got[:, ok], cent[:, lab[ok]])
    assert np.array_equal(got[:, ~ok], pts[:, ~ok])


def test_small_k_path_unchanged(torch_cuda):
    """k <= 1024 keeps the float fixed-order path: still within the bar."""
    torch = torch_cuda
    rng = np.random.default_rng(4)
    n, k = 100_000, 1024
    img = rng.random((3, n), dtype=np.float32)
    lab = rng.integers(0, k, size=n).astype(np.int32)
    C = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.calculate_centroids(_dev(torch, lab), _dev(torch, img), C)
    want, _ = _means_f64(lab, img, k)
    assert np.abs(C.cpu().numpy() - want).max() <= TOL
