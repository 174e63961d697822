"""CPU oracle of the warm-start network (numpy, float64 arithmetic).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the gfx950
network kernels.  The product path never imports it.

The network (no counterpart in the reference, whose cluster() takes its
initial centroids as an in/out argument, split/include/split/device/kmeans/
cluster.cuh:19-22,31-37 — this is what fills them):

    input  (B, 256, 256, 3) NHWC
    conv1  3x3 stride 2 pad 1,  3 -> 16, ReLU   -> (B, 128, 128, 16)
    conv2                      16 -> 32, ReLU   -> (B,  64,  64, 32)
    conv3                      32 -> 64, ReLU   -> (B,  32,  32, 64)
    conv4                      64 -> 64, ReLU   -> (B,  16,  16, 64)
    global average pool                          -> (B, 64)
    fc5    64 -> 128, ReLU
    fc6   128 -> 48, sigmoid  -> 16 centroids, value [c*16 + j] = channel c
                                  of centroid j (the (3, k) centroid layout)
    loss   (1/B) sum (p - t)^2 against luminance-sorted k-means centroids,
           or the Chamfer set distance between predicted and target colours

Parameters: one float32 vector; layer l holds W_l[n][0:Kp_l] row-major,
where columns 0..K_l-1 are the weights (conv: index (ky*3 + kx)*Cin + ci;
dense: ci), column K_l the bias, and the columns up to Kp_l (K_l + 1
rounded up to 32) zero.

With `mimic=True` every value the GPU kernels round to bfloat16 is rounded
here at the same point (weights, activations, the gradients fed to the
matrix cores); with `mimic=False` everything stays float64, which is what
the finite-difference check uses.
"""
from __future__ import annotations

import numpy as np

SIDE = 256
K_OUT = 16
# (kind, Cin, Cout)
LAYERS = [("conv", 3, 16), ("conv", 16, 32), ("conv", 32, 64), ("conv", 64, 64),
          ("dense", 64, 128), ("dense", 128, 3 * K_OUT)]


def layer_dims():
    """[(kind, cin, cout, K, Kp, offset)] and the total parameter count."""
    out, off = [], 0
    for kind, cin, cout in LAYERS:
        K = 9 * cin if kind == "conv" else cin
        Kp = (K + 1 + 31) // 32 * 32
        out.append((kind, cin, cout, K, Kp, off))
        off += cout * Kp
    return out, off


DIMS, PARAM_COUNT = layer_dims()


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed: int) -> np.ndarray:
    """He-normal weights, zero biases (the GPU trainer starts from these)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(PARAM_COUNT, dtype=np.float32)
    for kind, cin, cout, K, Kp, off in DIMS:
        w = rng.standard_normal((cout, K)) * np.sqrt(2.0 / K)
        blk = np.zeros((cout, Kp), dtype=np.float32)
        blk[:, :K] = w
        p[off:off + cout * Kp] = blk.ravel()
    return p


def weights(params, l, mimic):
    kind, cin, cout, K, Kp, off = DIMS[l]
    W = np.asarray(params[off:off + cout * Kp], dtype=np.float64).reshape(cout, Kp)
    return bf16(W) if mimic else W


def im2col(X):
    """(B, H, W, C) -> (B*Ho*Wo, 9C + 1) for 3x3 stride 2 pad 1, column
    (ky*3 + kx)*C + ci, and a trailing column of ones (the bias input)."""
    B, H, Wd, C = X.shape
    Ho, Wo = (H + 1) // 2, (Wd + 1) // 2
    P = np.zeros((B, H + 2, Wd + 2, C), dtype=np.float64)
    P[:, 1:H + 1, 1:Wd + 1, :] = X
    cols = np.empty((B, Ho, Wo, 9 * C + 1), dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            t = ky * 3 + kx
            cols[..., t * C:(t + 1) * C] = P[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :]
    cols[..., 9 * C] = 1.0
    return cols.reshape(B * Ho * Wo, 9 * C + 1)


def col2im(dcols, shape):
    """Adjoint of im2col (without the ones column): (M, 9C) -> (B, H, W, C)."""
    B, H, Wd, C = shape
    Ho, Wo = (H + 1) // 2, (Wd + 1) // 2
    d = dcols.reshape(B, Ho, Wo, 9 * C)
    P = np.zeros((B, H + 2, Wd + 2, C), dtype=np.float64)
    for ky in range(3):
        for kx in range(3):
            t = ky * 3 + kx
            P[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2, :] += d[..., t * C:(t + 1) * C]
    return P[:, 1:H + 1, 1:Wd + 1, :]


def forward(params, X0, mimic=True):
    """X0: (B, H, W, 3) (bf16 values when mimicking).  Returns the dict of
    activations the backward pass needs, with 'P' the (B, 48) outputs."""
    r = bf16 if mimic else (lambda v: np.asarray(v, dtype=np.float64))
    acts = {"X0": np.asarray(X0, dtype=np.float64)}
    X = acts["X0"]
    for l in range(4):
        kind, cin, cout, K, Kp, off = DIMS[l]
        B, H, Wd, _ = X.shape
        Ho, Wo = (H + 1) // 2, (Wd + 1) // 2
        cols = im2col(X)
        Z = cols @ weights(params, l, mimic)[:, :K + 1].T
        X = r(np.maximum(Z, 0.0)).reshape(B, Ho, Wo, cout)
        acts[f"Y{l + 1}"] = X
    B = X.shape[0]
    F = X.reshape(B, -1, X.shape[-1]).mean(axis=1)
    acts["F"] = F
    A5 = np.concatenate([r(F), np.ones((B, 1))], axis=1)
    H5 = r(np.maximum(A5 @ weights(params, 4, mimic)[:, :65].T, 0.0))
    acts["H"] = H5
    A6 = np.concatenate([H5, np.ones((B, 1))], axis=1)
    Z6 = A6 @ weights(params, 5, mimic)[:, :129].T
    acts["Z6"] = Z6
    acts["P"] = 1.0 / (1.0 + np.exp(-Z6))
    return acts


def loss(P, T, kind="mse"):
    """mse: (1/B) sum (p - t)^2 in slot order.  chamfer: (1/B) sum over
    images of sum_t min_p |t - p|^2 + sum_p min_t |p - t|^2 over the 16
    predicted and 16 target colours (a set distance: slot order does not
    matter)."""
    if kind == "mse":
        return float(((P - T) ** 2).sum() / P.shape[0])
    tot = 0.0
    for b in range(P.shape[0]):
        D = _pair_d2(P[b], T[b])
        tot += D.min(axis=0).sum() + D.min(axis=1).sum()
    return tot / P.shape[0]


def _pair_d2(p, t):
    """D[j, i] = |p_j - t_i|^2 for (48,) vectors in the (3, 16) layout."""
    pp = np.asarray(p, dtype=np.float64).reshape(3, K_OUT)
    tt = np.asarray(t, dtype=np.float64).reshape(3, K_OUT)
    return ((pp[:, :, None] - tt[:, None, :]) ** 2).sum(axis=0)


def chamfer_grad(P, T):
    """d loss(P, T, 'chamfer') / dP, ties to the lower index."""
    B = P.shape[0]
    G = np.zeros((B, 3, K_OUT))
    for b in range(B):
        pp = P[b].reshape(3, K_OUT)
        tt = T[b].reshape(3, K_OUT)
        D = _pair_d2(P[b], T[b])
        for i in range(K_OUT):       # each target pulls its nearest prediction
            j = int(np.argmin(D[:, i]))
            G[b, :, j] += 2.0 * (pp[:, j] - tt[:, i])
        for j in range(K_OUT):       # each prediction moves to its nearest target
            i = int(np.argmin(D[j, :]))
            G[b, :, j] += 2.0 * (pp[:, j] - tt[:, i])
    return G.reshape(B, 3 * K_OUT) / B


def backward(params, acts, T, mimic=True, kind="mse"):
    """Gradient of loss(P, T, kind) w.r.t. params (float64 vector, padded
    columns zero) and the loss."""
    r = bf16 if mimic else (lambda v: np.asarray(v, dtype=np.float64))
    P = acts["P"]
    B = P.shape[0]
    g = np.zeros(PARAM_COUNT, dtype=np.float64)

    def put(l, dW):
        _, cin, cout, K, Kp, off = DIMS[l]
        blk = np.zeros((cout, Kp))
        blk[:, :K + 1] = dW
        g[off:off + cout * Kp] = blk.ravel()

    dP = 2.0 * (P - T) / B if kind == "mse" else chamfer_grad(P, T)
    dZ6 = dP * P * (1.0 - P)
    G6 = r(dZ6)
    H5 = acts["H"]
    put(5, G6.T @ np.concatenate([H5, np.ones((B, 1))], axis=1))
    dH = G6 @ weights(params, 5, mimic)[:, :128]
    G5 = r(dH * (H5 > 0))
    put(4, G5.T @ np.concatenate([r(acts["F"]), np.ones((B, 1))], axis=1))
    dF = G5 @ weights(params, 4, mimic)[:, :64]
    Y4 = acts["Y4"]
    dY = np.broadcast_to((dF / (Y4.shape[1] * Y4.shape[2]))[:, None, None, :], Y4.shape)
    for l in range(3, -1, -1):
        _, cin, cout, K, Kp, off = DIMS[l]
        Y = acts[f"Y{l + 1}"]
        Xin = acts[f"Y{l}"] if l > 0 else acts["X0"]
        G = r(dY * (Y > 0)).reshape(-1, cout)
        put(l, G.T @ im2col(Xin))
        if l > 0:
            dcols = G @ weights(params, l, mimic)[:, :K]
            dY = col2im(dcols, Xin.shape)
    return g, loss(P, T, kind)


def sort_by_luminance(cent):
    """(3, k) centroids -> (3, k) ordered by Y = 0.299 R + 0.587 G + 0.114 B
    (float32 arithmetic), ties to the lower index: the training target
    layout."""
    c = np.asarray(cent, dtype=np.float32)
    y = (np.float32(0.299) * c[0] + np.float32(0.587) * c[1]) + np.float32(0.114) * c[2]
    order = sorted(range(c.shape[1]), key=lambda j: (y[j], j))
    return c[:, order]


def pack_input(planes, width, height, side=SIDE):
    """SoA float image -> (side, side, 3) area-average resample: output cell
    (X, Y) averages source columns [floor(X*W/side), max(floor((X+1)*W/side),
    x0 + 1)) and rows likewise (summed in float32, x inner, then divided)."""
    img = np.asarray(planes, dtype=np.float32).reshape(3, height, width)
    out = np.empty((side, side, 3), dtype=np.float32)
    xs = [(X * width // side, max((X + 1) * width // side, X * width // side + 1))
          for X in range(side)]
    ys = [(Y * height // side, max((Y + 1) * height // side, Y * height // side + 1))
          for Y in range(side)]
    for Y, (y0, y1) in enumerate(ys):
        for X, (x0, x1) in enumerate(xs):
            for c in range(3):
                s = np.float32(0.0)
                for yy in range(y0, y1):
                    for xx in range(x0, x1):
                        s = np.float32(s + img[c, yy, xx])
                out[Y, X, c] = s / np.float32((y1 - y0) * (x1 - x0))
    return out
