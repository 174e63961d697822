"""Colour conversions and the beta feature on the GPU, through the C ABI.

Bars (written here, float32 kernels against the FP64 oracle of
oracle/color.py): every conversion within a few float32 ulp of its output
scale — absolute 2e-4 for XYZ / Lab outputs (scale 100), 1e-5 for RGB
outputs (scale 1: the inverse companding's slope, about 13 near its
threshold, amplifies the float32 error of the linear value; measured up to
6.4e-6), relative 2e-6 for intensity-chroma.  In place equals out
of place bit for bit.  beta is bit-exact against the oracle's float32
thresholds applied to the GPU's own Lab planes, and equal to the FP64 chain
except for pixels within 1e-3 of a threshold; the fused sequence equals the
separate calls bit for bit.
"""
import numpy as np
import pytest

import oracle.color as OC
import split_amd as sa
from split_amd import workloads as W
from tests.conftest import say

pytestmark = pytest.mark.gpu

ATOL = {"rgb_to_xyz": 2e-4, "xyz_to_rgb": 1e-5, "xyz_to_lab": 2e-4, "lab_to_xyz": 2e-4,
        "rgb_to_lab": 2e-4, "lab_to_rgb": 1e-5}
IC_RTOL = 2e-6


def _inputs(conv, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = (0.05, 1.0) if conv == "rgb_to_ic" else (-0.1, 1.2)
    rgb = rng.uniform(lo, hi, (3, n))
    if conv not in ("rgb_to_ic",):
        # companding thresholds and the ends of the range, exactly
        edge = np.array([0.0, 0.04045, 1.0, 0.5, 0.0031308])
        rgb[:, :min(n, 5)] = edge[:min(n, 5)]
    rgb = rgb.astype(np.float32)
    src = conv.split("_to_")[0]
    if src == "rgb":
        return rgb
    if src == "xyz":
        return OC.rgb_to_xyz(np.clip(rgb, 0, 1)).astype(np.float32)
    return OC.rgb_to_lab(np.clip(rgb, 0, 1)).astype(np.float32)


@pytest.mark.parametrize("conv", sorted(sa.CONVERSIONS))
@pytest.mark.parametrize("n", [1, 7, 4096, 262_147])
def test_conversion_matches_fp64(torch_cuda, conv, n):
    torch = torch_cuda
    say(f"color: {conv} n={n}")
    x = _inputs(conv, n, seed=n)
    X = torch.from_numpy(x).cuda()
    Y = torch.empty_like(X)
    sa.convert_color_space(X, Y, conv)
    torch.cuda.synchronize()
    got = Y.cpu().numpy().astype(np.float64)
    want = OC.CONVERT[conv](x.astype(np.float64))
    if conv == "rgb_to_ic":
        err = np.abs(got - want) / np.abs(want)
        assert err.max() <= IC_RTOL, f"{conv}: max rel err {err.max():.3g}"
    else:
        err = np.abs(got - want).max()
        assert err <= ATOL[conv], f"{conv}: max abs err {err:.3g}"
    # in place gives the same bits
    Z = X.clone()
    sa.convert_color_space(Z, Z, conv)
    torch.cuda.synchronize()
    assert torch.equal(Z, Y)


def test_conversion_unaligned_planes(torch_cuda):
    torch = torch_cuda
    say("color: unaligned planes")
    n = 1001
    x = _inputs("rgb_to_lab", n, seed=5)
    buf = torch.zeros(3 * n + 1, dtype=torch.float32, device="cuda")
    X = buf[1:].view(3, n)
    X.copy_(torch.from_numpy(x))
    Y = torch.empty((3, n), dtype=torch.float32, device="cuda")
    sa.convert_color_space(X, Y, "rgb_to_lab")
    torch.cuda.synchronize()
    assert np.abs(Y.cpu().numpy() - OC.rgb_to_lab(x.astype(np.float64))).max() <= 2e-4


def test_conversion_errors(torch_cuda):
    torch = torch_cuda
    say("color: error paths")
    X = torch.rand((3, 64), device="cuda")
    L = sa.lib()
    assert L.split_hip_convert_color_space(X.data_ptr(), X.data_ptr(), 64, 99, None) == 1
    # out overlapping in by one pixel (not equal): rejected
    buf = torch.rand(3 * 64 + 1, device="cuda")
    assert L.split_hip_convert_color_space(buf.data_ptr(), buf.data_ptr() + 4, 64, 0,
                                           None) == 1
    assert L.split_hip_convert_color_space(None, X.data_ptr(), 64, 0, None) == 1


def _feature_image(side=256, seed=9):
    """A shaded Voronoi image plus blocks of dark and bright grey, so that
    both non-zero beta values occur."""
    w = W.scaled(W.PLAIN_512, side, side)
    img = W.make_image(w.image).reshape(3, side, side).copy()
    rng = np.random.default_rng(seed)
    img[:, :16, :16] = rng.uniform(0.0, 0.03, (16, 16))  # dark grey
    img[:, -16:, -16:] = rng.uniform(0.985, 1.0, (16, 16))  # bright grey
    return np.ascontiguousarray(img.reshape(3, side * side).astype(np.float32))


def test_beta_feature_exact(torch_cuda):
    torch = torch_cuda
    say("color: beta_feature")
    rgb = _feature_image()
    n = rgb.shape[1]
    P = torch.from_numpy(rgb).cuda()
    lab = torch.empty_like(P)
    sa.convert_color_space(P, lab, "rgb_to_lab")
    beta = torch.empty(n, dtype=torch.float32, device="cuda")
    sa.beta_feature(lab, beta, mu=1e5)
    torch.cuda.synchronize()
    lab_h = lab.cpu().numpy()
    got = beta.cpu().numpy()
    want = OC.beta_feature(lab_h, 1e5)
    assert np.array_equal(got, want)
    assert (got > 0).sum() >= 200 and (got < 0).sum() >= 200, "both signs must occur"
    # against the FP64 chain, away from the thresholds
    lab64 = OC.rgb_to_lab(rgb.astype(np.float64))
    want64 = OC.beta_feature(lab64.astype(np.float32), 1e5)
    far = OC.beta_margin(lab64) > 1e-3
    assert far.mean() > 0.99
    assert np.array_equal(got[far], want64[far])
    # in place over the L plane (the sample's call)
    lab2 = lab.clone()
    sa.beta_feature(lab2, lab2[0], mu=1e5)
    torch.cuda.synchronize()
    assert torch.equal(lab2[0], beta) and torch.equal(lab2[1:], lab[1:])


@pytest.mark.parametrize("n", [65536, 65535, 3])
def test_lab_beta_fused_equals_sequence(torch_cuda, n):
    torch = torch_cuda
    say(f"color: fused lab+beta n={n}")
    rgb = np.ascontiguousarray(_feature_image()[:, :n])
    P = torch.from_numpy(rgb).cuda()
    lab = torch.empty_like(P)
    sa.convert_color_space(P, lab, "rgb_to_lab")
    beta = torch.empty(n, dtype=torch.float32, device="cuda")
    sa.beta_feature(lab, beta, mu=3.0)
    F = torch.empty_like(P)
    lum = torch.empty(n, dtype=torch.float32, device="cuda")
    sa.lab_beta_feature(P, F, lum, mu=3.0)
    torch.cuda.synchronize()
    assert torch.equal(lum, lab[0])
    assert torch.equal(F[1:], lab[1:])
    assert torch.equal(F[0], beta)
    # in place over the RGB planes, no luminance copy
    Q = P.clone()
    sa.lab_beta_feature(Q, Q, None, mu=3.0)
    torch.cuda.synchronize()
    assert torch.equal(Q, F)


def test_beta_errors(torch_cuda):
    torch = torch_cuda
    say("color: beta error paths")
    L = sa.lib()
    lab = torch.rand((3, 64), device="cuda")
    beta = torch.empty(64, device="cuda")
    t = torch.empty(16, dtype=torch.uint8, device="cuda")
    assert L.split_hip_beta_feature(lab.data_ptr(), 64, beta.data_ptr(), 1.0, t.data_ptr(),
                                    16, None) == 2
    # beta straddling two planes of lab
    t2 = torch.empty(4096, dtype=torch.uint8, device="cuda")
    assert L.split_hip_beta_feature(lab.data_ptr(), 64, lab.data_ptr() + 4 * 32, 1.0,
                                    t2.data_ptr(), 4096, None) == 1
