"""split_hip_calculate_centroids_noalloc through the C ABI: the allocation-
free variant (static per-GPU scratch) gives exactly what
split_hip_calculate_centroids gives with its own workspace, on two calls in
a row, the second with a smaller input (the scratch keeps the first call's
contents, which must not leak into the second).

Bar: bit-identical centroids.  For k <= 1024 both run the same kernels on
the same grid (float sums in the same fixed order); for larger k both sum
quantised integers, which do not depend on order.  Sizes beyond the static
scratch are refused on the host, before any kernel."""
import numpy as np
import pytest

import oracle
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _both(torch, labels, pts, k):
    """(existing entry point with its workspace, no-alloc variant)."""
    want = torch.full((3, k), -3.0, dtype=torch.float32, device="cuda")
    sa.calculate_centroids(labels, pts, want)
    got = torch.full((3, k), -5.0, dtype=torch.float32, device="cuda")
    sa.calculate_centroids_noalloc(labels, pts, got)
    torch.cuda.synchronize()
    return want.cpu().numpy(), got.cpu().numpy()


def _labelled(torch, spec, k, seed):
    img = W.make_image(spec)
    P = _dev(torch, img)
    C = torch.empty((3, k), dtype=torch.float32, device="cuda")
    sa.uniform_random_initialize(P, C, seed)
    L = torch.empty(img.shape[1], dtype=torch.int32, device="cuda")
    sa.label_points(C, P, L, width=spec.width, height=spec.height)
    return P, L


@pytest.mark.parametrize("k", [8, 32, 256, 1024])
def test_noalloc_equals_workspace_call(torch_cuda, k):
    torch = torch_cuda
    big = W.ImageSpec(2048, 1024, seed=0xA110C + k, sites=300, shading=True,
                      noise_sigma=0.01)
    small = W.ImageSpec(333, 517, seed=0xA110D + k, sites=40, shading=True,
                        noise_sigma=0.01)
    for tag, spec in (("first", big), ("second, smaller", small)):
        say(f"noalloc calculate_centroids k={k}: {tag} call {spec.width}x{spec.height}")
        P, L = _labelled(torch, spec, k, 0x5EED + k)
        want, got = _both(torch, L, P, k)
        assert np.array_equal(want, got), tag
        # and both are the FP64 means
        m, _ = oracle.centroids_f64(L.cpu().numpy(), P.cpu().numpy(), k)
        assert np.abs(got - m).max() <= 1e-5, tag


def test_noalloc_segment_means(torch_cuda):
    """k = number of segments (> 1024): the segment-mean kernels in the
    static scratch, then a smaller image with fewer segments."""
    torch = torch_cuda
    for tag, (w, h, sites) in (("first", (1024, 1024, 400)), ("second, smaller", (512, 300, 90))):
        spec = W.ImageSpec(w, h, seed=0x5E65 + w, sites=sites, shading=True, noise_sigma=0.02)
        P = _dev(torch, W.make_image(spec))
        C = torch.empty((3, 8), dtype=torch.float32, device="cuda")
        sa.uniform_random_initialize(P, C, 0x77 + w)
        L = torch.empty(w * h, dtype=torch.int32, device="cuda")
        S = torch.empty(w * h, dtype=torch.int32, device="cuda")
        r = sa.segment(P, w, h, C, L, S, max_iter=10, threshold=-1.0, connectivity=4)
        k = r.num_segments
        say(f"noalloc segment means: {tag} call {w}x{h}, k={k}")
        assert k > 1024
        want, got = _both(torch, S, P, k)
        assert np.array_equal(want, got), tag
        m, _ = oracle.centroids_f64(S.cpu().numpy(), P.cpu().numpy(), k)
        assert np.abs(got - m).max() <= 1e-5, tag


def test_noalloc_refuses_beyond_scratch(torch_cuda):
    torch = torch_cuda
    say("noalloc calculate_centroids: k beyond the static scratch is refused")
    P = torch.zeros((3, 16), dtype=torch.float32, device="cuda")
    L = torch.zeros(16, dtype=torch.int32, device="cuda")
    C = torch.zeros((3, 700_000), dtype=torch.float32, device="cuda")
    with pytest.raises(sa.SplitError) as e:
        sa.calculate_centroids_noalloc(L, P, C)
    assert e.value.status == 5  # UNSUPPORTED
    assert "static scratch" in str(e.value)
