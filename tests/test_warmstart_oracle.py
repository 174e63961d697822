"""Pins of the warm-start network oracle (CPU only).

The oracle's forward pass is checked against a direct 7-loop convolution,
its im2col/col2im pair against the adjoint identity, its whole backward pass
against central finite differences of the float64 loss, and its helpers
against closed forms.
"""
import numpy as np
import pytest

from oracle import warmstart as ows


def _conv_direct(X, W, b):
    """3x3 stride-2 pad-1 convolution by definition; W (Cout, 3, 3, Cin)."""
    B, H, Wd, C = X.shape
    Ho, Wo = (H + 1) // 2, (Wd + 1) // 2
    out = np.zeros((B, Ho, Wo, W.shape[0]))
    for bb in range(B):
        for oy in range(Ho):
            for ox in range(Wo):
                for co in range(W.shape[0]):
                    s = b[co]
                    for ky in range(3):
                        for kx in range(3):
                            iy, ix = 2 * oy + ky - 1, 2 * ox + kx - 1
                            if 0 <= iy < H and 0 <= ix < Wd:
                                s += np.dot(W[co, ky, kx], X[bb, iy, ix])
                    out[bb, oy, ox, co] = s
    return out


def test_im2col_matches_direct_convolution():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2, 7, 6, 3))
    W = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    Wm = np.concatenate([W.reshape(4, 27), b[:, None]], axis=1)
    got = (ows.im2col(X) @ Wm.T).reshape(2, 4, 3, 4)
    assert np.allclose(got, _conv_direct(X, W, b), rtol=0, atol=1e-12)


def test_col2im_is_adjoint_of_im2col():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((2, 9, 8, 5))
    D = rng.standard_normal((2 * 5 * 4, 45))
    lhs = np.sum(ows.im2col(X)[:, :45] * D)
    rhs = np.sum(X * ows.col2im(D, X.shape))
    assert abs(lhs - rhs) < 1e-9 * max(1.0, abs(lhs))


@pytest.mark.parametrize("kind", ["mse", "chamfer"])
def test_backward_matches_finite_differences(kind):
    rng = np.random.default_rng(3)
    params = ows.init_params(7).astype(np.float64)
    params += 0.05 * rng.standard_normal(params.shape) * (params != 0)
    # biases away from zero so ReLUs are not all at their kink
    for kind, cin, cout, K, Kp, off in ows.DIMS:
        blk = params[off:off + cout * Kp].reshape(cout, Kp)
        blk[:, K] = 0.05 * rng.standard_normal(cout)
    X = rng.random((2, 16, 16, 3))
    T = rng.random((2, 48))
    acts = ows.forward(params, X, mimic=False)
    g, L = ows.backward(params, acts, T, mimic=False, kind=kind)
    assert abs(L - ows.loss(acts["P"], T, kind)) <= 1e-12 * L
    eps = 1e-6
    checked = 0
    for l, (kind, cin, cout, K, Kp, off) in enumerate(ows.DIMS):
        idx = [off + n * Kp + k for n, k in zip(rng.integers(0, cout, 6),
                                                rng.integers(0, K + 1, 6))]
        idx.append(off + (cout - 1) * Kp + K)  # a bias
        for i in idx:
            p1 = params.copy()
            p1[i] += eps
            p0 = params.copy()
            p0[i] -= eps
            fd = (ows.loss(ows.forward(p1, X, False)["P"], T, kind) -
                  ows.loss(ows.forward(p0, X, False)["P"], T, kind)) / (2 * eps)
            assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-10, (l, i, fd, g[i])
            checked += 1
        # padded columns have zero gradient
        for k in range(K + 1, Kp):
            assert g[off + k] == 0.0
    assert checked == 6 * 7


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -2.5, 0.0],
                 dtype=np.float32)
    got = ows.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: ties to even -> 1
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2.0 ** -7
    assert got[3] == -2.5 and got[4] == 0.0


def test_sort_by_luminance():
    c = np.array([[0.9, 0.1, 0.5, 0.5],
                  [0.9, 0.1, 0.5, 0.5],
                  [0.9, 0.1, 0.5, 0.5]], dtype=np.float32)
    s = ows.sort_by_luminance(c)
    assert np.array_equal(s[0], np.array([0.1, 0.5, 0.5, 0.9], dtype=np.float32))


def test_pack_input_identity_and_box_mean():
    rng = np.random.default_rng(4)
    img = rng.random((3, 256 * 256), dtype=np.float32)
    out = ows.pack_input(img, 256, 256)
    assert np.array_equal(out, img.reshape(3, 256, 256).transpose(1, 2, 0))
    const = np.full((3, 300 * 270), 0.25, dtype=np.float32)
    assert np.all(ows.pack_input(const, 300, 270) == np.float32(0.25))


@pytest.mark.parametrize("B", [1, 3])
def test_forward_shapes_and_range(B):
    params = ows.init_params(5)
    X = np.random.default_rng(6).random((B, 256, 256, 3))
    acts = ows.forward(params, ows.bf16(X), mimic=True)
    assert acts["Y1"].shape == (B, 128, 128, 16)
    assert acts["Y4"].shape == (B, 16, 16, 64)
    assert acts["P"].shape == (B, 48)
    assert np.all((acts["P"] > 0) & (acts["P"] < 1))


def test_chamfer_closed_forms():
    rng = np.random.default_rng(11)
    T = rng.random((2, 48))
    # identical sets in any slot order: zero loss and zero gradient
    perm = rng.permutation(16)
    P = T.reshape(2, 3, 16)[:, :, perm].reshape(2, 48)
    assert ows.loss(P, T, "chamfer") == 0.0
    assert np.all(ows.chamfer_grad(P, T) == 0.0)
    # all predictions at one point c: loss = sum_t |t-c|^2 + 16 min_t |c-t|^2
    c = np.array([0.2, 0.5, 0.7])
    P1 = np.repeat(c, 16)[None, :]
    d = ((T[0].reshape(3, 16) - c[:, None]) ** 2).sum(0)
    assert abs(ows.loss(P1, T[:1], "chamfer") - (d.sum() + 16 * d.min())) < 1e-12
