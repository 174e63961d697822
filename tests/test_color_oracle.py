"""Pins of the colour / transfer oracle (oracle/color.py) on the CPU.

Forward conversions against published values (sRGB primaries in CIE Lab,
the D65 white point), inverse conversions by round trips and the product of
the two matrices, companding branches by continuity, beta_feature by hand-
built cases (including the reference's lambda from the a plane only), and
quantisation by known answers.
"""
import numpy as np
import pytest

import oracle.color as C


def test_srgb_primaries_lab_published():
    # sRGB primaries, white, black in CIE Lab (D65), as published to 4
    # decimals (e.g. Lindbloom's calculator / OpenCV's float Lab)
    rgb = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [0, 0, 0]], float).T
    want = np.array([[53.2408, 80.0925, 67.2032],
                     [87.7347, -86.1827, 83.1793],
                     [32.2970, 79.1875, -107.8602],
                     [100.0, 0.0, 0.0],
                     [0.0, 0.0, 0.0]]).T
    assert np.abs(C.rgb_to_lab(rgb) - want).max() < 6e-4


def test_white_is_d65():
    xyz = C.rgb_to_xyz(np.ones((3, 1)))[:, 0]
    assert np.abs(xyz - [95.047, 100.0, 108.883]).max() < 1e-4


def test_matrices_are_inverse():
    assert np.abs(C.M_XYZ_RGB @ C.M_RGB_XYZ - np.eye(3)).max() < 3e-7


@pytest.mark.parametrize("lo,hi", [(0.0, 1.0), (0.0, 0.05)])
def test_round_trips(lo, hi):
    rng = np.random.default_rng(3)
    rgb = rng.uniform(lo, hi, (3, 2000))
    assert np.abs(C.xyz_to_rgb(C.rgb_to_xyz(rgb)) - rgb).max() < 3e-6
    assert np.abs(C.lab_to_rgb(C.rgb_to_lab(rgb)) - rgb).max() < 3e-6
    xyz = C.rgb_to_xyz(rgb)
    assert np.abs(C.lab_to_xyz(C.xyz_to_lab(xyz)) - xyz).max() < 1e-9


def test_grey_has_no_chroma():
    v = np.linspace(0, 1, 11)
    lab = C.rgb_to_lab(np.stack([v, v, v]))
    assert np.abs(lab[1:]).max() < 2e-3  # matrix constants to 7 digits
    assert np.all(np.diff(lab[0]) > 0)


def test_companding_branches_meet():
    t = 0.04045
    assert abs(((t + 0.055) / 1.055) ** 2.4 - t / 12.92) < 1e-7
    e = C.EPS
    assert abs(np.cbrt(e) - (C.KAPPA * e + 16 / 116)) < 1e-6


def test_rgb_to_ic_closed_form():
    ic = C.rgb_to_ic(np.array([[0.2], [0.4], [0.6]]))[:, 0]
    assert np.allclose(ic, [0.4, 0.5, 1.0])


def _lab(cols):
    return np.array(cols, dtype=np.float32).T


def test_beta_hand_cases():
    # max L = 100 -> lo 20, hi 95; max |a| = 50 -> lambda 10
    lab = _lab([[100, 0, 0],      # bright grey -> +mu
                [10, 5, -5],      # dark grey -> -mu
                [50, 1, 1],       # mid grey -> 0
                [10, 50, 0],      # dark but chromatic (a) -> 0
                [10, 0, 12],      # dark but chromatic (b) -> 0
                [20, 0, 0],       # L == lo (strict) -> 0
                [95, 0, 0],       # L == hi (strict) -> 0
                [10, 10, 0],      # |a| == lambda (strict) -> 0
                [96, -9, 9]])     # bright, |a|,|b| < lambda -> +mu
    b = C.beta_feature(lab, mu=7.0)
    assert b.tolist() == [7, -7, 0, 0, 0, 0, 0, 0, 7]


def test_beta_lambda_from_a_plane_only():
    # max |b| = 200 would give lambda 40; the reference takes max |a| = 20
    # -> lambda 4, so |b| = 30 is not grey
    lab = _lab([[100, 20, 0], [5, 0, 30], [5, 0, 3], [5, 0, -200]])
    assert C.beta_feature(lab, mu=1.0).tolist() == [0, 0, -1, 0]


def test_quantize_known_answers():
    v = np.array([[0.0, 1.0, 0.5, -1.0, 2.0, 0.25, 0.1]], np.float32)
    # 0.25*255+0.5 = 64.25 -> 64; 0.1*255+0.5 = 26.0000004 -> 26
    assert C.quantize(v)[:, 0].tolist() == [0, 255, 128, 0, 255, 64, 26]


def test_quantize_single_rounding():
    # v*255 + 0.5 just below an integer in exact arithmetic rounds up to it
    # when the multiply-add is rounded once to float32
    v = np.float32(np.nextafter(np.float32(0.5) / np.float32(255), np.float32(0)))
    exact = float(v) * 255 + 0.5
    assert exact < 1.0
    assert C.quantize(np.array([[v]], np.float32))[0, 0] == (1 if np.float32(exact) == 1 else 0)


def test_layout_inverse():
    rng = np.random.default_rng(0)
    for c in (1, 2, 3, 4):
        planes = rng.random((c, 37)).astype(np.float32)
        hwc = C.interleave(planes)
        assert hwc.shape == (37, c)
        assert np.array_equal(hwc[5], planes[:, 5])
        assert np.array_equal(C.deinterleave(hwc.ravel(), c), planes)
