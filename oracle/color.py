"""CPU oracle of the colour features and the image transfer formats (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and bench.py's CPU leg) as the
checker of the gfx950 colour, layout and quantisation kernels.  The product
path never imports it.

Conversions are restated in float64 from the reference's formulas
(split/include/split/device/color/detail/conversion.inl), which follow the
sRGB / CIE definitions with a D65 white point:

    rgb_to_xyz   conversion.inl:13-31   sRGB companding, x100, sRGB->XYZ matrix
    xyz_to_rgb   conversion.inl:33-51   XYZ->sRGB matrix, inverse companding
    xyz_to_lab   conversion.inl:53-72   white (95.047, 100, 108.883), f(t) with
                                        the 0.008856 / 7.787 approximations
    lab_to_xyz   conversion.inl:74-94   the inverse of f
    rgb_to_lab   conversion.inl:96-100, lab_to_rgb conversion.inl:102-106
    rgb_to_ic    conversion.inl:108-117 intensity, r / intensity, g / intensity

beta_feature follows split/src/split/device/color/beta_feature.cu:15-78
(lambda from the a plane only, thresholds 0.2 and 0.95 of max L, strict
comparisons), computed in float32 like the reference's host-side threshold
arithmetic so that it can be compared bit for bit.

Arrays are SoA: (3, N) planes.
"""
from __future__ import annotations

import numpy as np

# sRGB (linear) -> XYZ and XYZ -> sRGB matrices, D65 (conversion.inl:27-30,
# 41-43); the published sRGB constants.
M_RGB_XYZ = np.array([[0.4124564, 0.3575761, 0.1804375],
                      [0.2126729, 0.7151522, 0.0721750],
                      [0.0193339, 0.1191920, 0.9503041]])
M_XYZ_RGB = np.array([[3.2404542, -1.5371385, -0.4985314],
                      [-0.9692660, 1.8760108, 0.0415560],
                      [0.0556434, -0.2040259, 1.0572252]])
WHITE = np.array([95.047, 100.0, 108.883])
EPS = 0.008856
KAPPA = 7.787


def _f64(p):
    return np.asarray(p, dtype=np.float64)


def rgb_to_xyz(rgb):
    """conversion.inl:13-31"""
    v = _f64(rgb)
    lin = np.where(v > 0.04045, ((np.maximum(v, 0.04045) + 0.055) / 1.055) ** 2.4,
                   v / 12.92) * 100.0
    return M_RGB_XYZ @ lin


def xyz_to_rgb(xyz):
    """conversion.inl:33-51"""
    lin = M_XYZ_RGB @ (_f64(xyz) / 100.0)
    return np.where(lin > 0.0031308,
                    1.055 * np.maximum(lin, 0.0031308) ** (1.0 / 2.4) - 0.055,
                    12.92 * lin)


def _lab_f(t):
    return np.where(t > EPS, np.cbrt(t), KAPPA * t + 16.0 / 116.0)


def _lab_finv(t):
    t3 = t * t * t
    return np.where(t3 > EPS, t3, (t - 16.0 / 116.0) / KAPPA)


def xyz_to_lab(xyz):
    """conversion.inl:53-72"""
    v = _f64(xyz) / WHITE[:, None]
    fx, fy, fz = _lab_f(v[0]), _lab_f(v[1]), _lab_f(v[2])
    return np.stack([116.0 * fy - 16.0, 500.0 * (fx - fy), 200.0 * (fy - fz)])


def lab_to_xyz(lab):
    """conversion.inl:74-94"""
    v = _f64(lab)
    fy = (v[0] + 16.0) / 116.0
    fx = v[1] / 500.0 + fy
    fz = fy - v[2] / 200.0
    return np.stack([_lab_finv(fx), _lab_finv(fy), _lab_finv(fz)]) * WHITE[:, None]


def rgb_to_lab(rgb):
    """conversion.inl:96-100"""
    return xyz_to_lab(rgb_to_xyz(rgb))


def lab_to_rgb(lab):
    """conversion.inl:102-106"""
    return xyz_to_rgb(lab_to_xyz(lab))


def rgb_to_ic(rgb):
    """conversion.inl:108-117"""
    v = _f64(rgb)
    i = (v[0] + v[1] + v[2]) / 3.0
    return np.stack([i, v[0] / i, v[1] / i])


CONVERT = {"rgb_to_xyz": rgb_to_xyz, "xyz_to_rgb": xyz_to_rgb, "xyz_to_lab": xyz_to_lab,
           "lab_to_xyz": lab_to_xyz, "rgb_to_lab": rgb_to_lab, "lab_to_rgb": lab_to_rgb,
           "rgb_to_ic": rgb_to_ic}


def beta_thresholds(lab):
    """(lambda, lo, hi) of beta_feature.cu:56-69 from float32 Lab planes, in
    float32 arithmetic: lambda = 0.2 * max |a|, lo = 0.2 * max L,
    hi = 0.95 * max L."""
    lab = np.asarray(lab, dtype=np.float32)
    lmax = np.float32(lab[0].max())
    amax = np.float32(np.abs(lab[1]).max())
    return (np.float32(0.2) * amax, lmax * np.float32(0.20), lmax * np.float32(0.95))


def beta_feature(lab, mu=1e5):
    """beta_feature.cu:15-45 (per pixel) with the thresholds of :56-69."""
    lab = np.asarray(lab, dtype=np.float32)
    lam, lo, hi = beta_thresholds(lab)
    L, a, b = lab[0], np.abs(lab[1]), np.abs(lab[2])
    grey = (a < lam) & (b < lam)
    mu = np.float32(mu)
    out = np.where(grey & (L < lo), -mu, np.where(grey & (L > hi), mu, np.float32(0)))
    return out.astype(np.float32)


def beta_margin(lab):
    """Per pixel, the smallest distance of |a|, |b| to lambda and of L to lo
    and hi: pixels closer than float rounding to a threshold may fall either
    way when lab is computed in another precision."""
    lab = np.asarray(lab, dtype=np.float64)
    lam, lo, hi = (float(x) for x in beta_thresholds(lab.astype(np.float32)))
    return np.minimum.reduce([np.abs(np.abs(lab[1]) - lam), np.abs(np.abs(lab[2]) - lam),
                              np.abs(lab[0] - lo), np.abs(lab[0] - hi)])


def lab_beta_feature(rgb, mu=1e5):
    """The sample's sequence (samples/src/main.cu:216-225) in float64 up to
    the Lab planes: returns (features with beta over L, luminance)."""
    lab = rgb_to_lab(rgb)
    feat = lab.copy()
    feat[0] = beta_feature(lab.astype(np.float32), mu)
    return feat, lab[0].copy()


# ---------------------------------------------------------------- layout ---

def deinterleave(hwc, channels):
    """(N*C,) or (..., C) interleaved -> (C, N) planes (samples/src/main.cu:98-114)."""
    a = np.asarray(hwc).reshape(-1, channels)
    return np.ascontiguousarray(a.T)


def interleave(planes):
    """(C, N) planes -> (N, C) interleaved (samples/src/main.cu:116-132)."""
    return np.ascontiguousarray(np.asarray(planes).T)


def quantize(planes):
    """stbi::quantize (split/src/split/host/stbi/stbi_raii.cpp:135-143) of
    (C, N) float32 planes into (N, C) uint8: clamp(v, 0, 1) * 255 + 0.5,
    the multiply-add rounded once to float32 (float64 holds the exact
    product-sum of a float32 and 255), then truncated."""
    v = np.clip(np.asarray(planes, dtype=np.float32), 0.0, 1.0).astype(np.float64)
    r = (v * 255.0 + 0.5).astype(np.float32)
    return np.ascontiguousarray(np.trunc(r).astype(np.uint8).T)
