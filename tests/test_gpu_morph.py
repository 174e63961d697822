"""Erosion on the GPU, through the C ABI, bit-exact against the restatement
of the reference's algorithm (oracle/morph.py) and its window definition."""
import numpy as np
import pytest

import oracle.morph as M
import split_amd as sa
from tests.conftest import say
from tests.test_morph_oracle import _blocky

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("h,w", [(1, 1), (1, 77), (77, 1), (31, 33), (257, 129), (512, 512)])
@pytest.mark.parametrize("n", [1, 2, 10, 64, 65])
def test_erode_matches_reference(torch_cuda, h, w, n):
    torch = torch_cuda
    say(f"morph: erode {h}x{w} n={n}")
    lab = _blocky(h, w, seed=h * 7 + w + n, values=3, block=max(2, 2 * n // 3 + 1))
    want = M.ref_erode(lab, n)
    L = torch.from_numpy(lab).cuda()
    sa.erode(L, n)
    torch.cuda.synchronize()
    assert np.array_equal(L.cpu().numpy(), want)


@pytest.mark.parametrize("h,w,n", [(1, 300, 3), (200, 130, 1), (129, 513, 70)])
def test_horizontal_sweep_matches_reference(torch_cuda, h, w, n):
    torch = torch_cuda
    say(f"morph: sweep {h}x{w} n={n}")
    lab = _blocky(h, w, seed=n, values=4, block=6)
    L = torch.from_numpy(lab).cuda()
    sa.erode_horizontal_sweep(L, n)
    torch.cuda.synchronize()
    assert np.array_equal(L.cpu().numpy(), M.ref_erode_horizontal_sweep(lab, n))


def test_erode_segment_labels(torch_cuda):
    """The sample's use: erosion of compressed segment labels (n = 10)."""
    torch = torch_cuda
    from split_amd import workloads as W
    say("morph: erode segment labels of a 512x384 shaded image")
    w = W.scaled(W.SHADED_4096, 512, 384)
    img = W.make_image(w.image)
    P = torch.from_numpy(img).cuda()
    C = torch.empty((3, w.k), dtype=torch.float32, device="cuda")
    sa.uniform_random_initialize(P, C, w.init_seed)
    Lb = torch.empty(w.n, dtype=torch.int32, device="cuda")
    S = torch.empty(w.n, dtype=torch.int32, device="cuda")
    sa.segment(P, 512, 384, C, Lb, S, max_iter=20, threshold=-1.0, connectivity=8)
    seg = S.cpu().numpy().reshape(384, 512)
    E = S.view(384, 512).clone()
    sa.erode(E, 10)
    torch.cuda.synchronize()
    got = E.cpu().numpy()
    assert np.array_equal(got, M.ref_erode(seg, 10))
    assert (got >= 0).any() and (got == -1).any()


def test_erode_zero_is_identity_and_errors(torch_cuda):
    torch = torch_cuda
    say("morph: n = 0 and error paths")
    lab = _blocky(20, 30, seed=1)
    L = torch.from_numpy(lab).cuda()
    sa.erode(L, 0)
    torch.cuda.synchronize()
    assert np.array_equal(L.cpu().numpy(), lab)
    lib = sa.lib()
    t = torch.empty(16, dtype=torch.uint8, device="cuda")
    assert lib.split_hip_erode(L.data_ptr(), 30, 20, -1, t.data_ptr(), 16, None) == 1
    assert lib.split_hip_erode(L.data_ptr(), 30, 20, 2, t.data_ptr(), 16, None) == 2
