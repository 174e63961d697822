"""Pins of the erosion oracle (oracle/morph.py) on the CPU: the restatement
of the reference's run-parity algorithm against the window definition on
random images, and closed forms (uniform image, checkerboard, stripes)."""
import numpy as np
import pytest

import oracle.morph as M


def _blocky(h, w, seed, values=3, block=4):
    rng = np.random.default_rng(seed)
    small = rng.integers(0, values, size=((h + block - 1) // block, (w + block - 1) // block))
    lab = np.kron(small, np.ones((block, block), dtype=np.int64))[:h, :w]
    noise = rng.random((h, w)) < 0.05
    lab[noise] = rng.integers(0, values, size=noise.sum())
    return lab.astype(np.int32)


@pytest.mark.parametrize("h,w", [(1, 1), (1, 17), (17, 1), (13, 29), (40, 33)])
@pytest.mark.parametrize("n", [1, 2, 3, 7])
def test_sweep_restatement_equals_window_definition(h, w, n):
    lab = _blocky(h, w, seed=h * 100 + w + n)
    assert np.array_equal(M.ref_erode_horizontal_sweep(lab, n), M.erode_window(lab, n, 1))


@pytest.mark.parametrize("n", [1, 2, 5])
def test_erode_is_square_window(n):
    lab = _blocky(37, 41, seed=n, values=2, block=6)
    got = M.ref_erode(lab, n)
    h, w = lab.shape
    want = np.full((h, w), -1, np.int32)
    for r in range(n, h - n):
        for c in range(n, w - n):
            sq = lab[r - n:r + n + 1, c - n:c + n + 1]
            if np.all(sq == lab[r, c]):
                want[r, c] = lab[r, c]
    assert np.array_equal(got, want)
    # columns first, then rows, as two window sweeps
    assert np.array_equal(got, M.erode_window(M.erode_window(lab, n, 0), n, 1))


def test_closed_forms():
    u = np.full((20, 30), 7, np.int32)
    e = M.ref_erode(u, 3)
    assert (e[3:17, 3:27] == 7).all() and (e == 7).sum() == 14 * 24
    cb = (np.indices((16, 16)).sum(0) % 2).astype(np.int32)
    assert (M.ref_erode(cb, 1) == -1).all()
    stripes = np.repeat(np.arange(5), 8)[None, :].repeat(4, 0).astype(np.int32)
    s = M.ref_erode_horizontal_sweep(stripes, 2)
    assert (s[:, 2:6] == 0).all() and (s[:, 0:2] == -1).all() and (s[:, 6:10] == -1).all()


def test_reference_undefined_for_zero():
    with pytest.raises(ValueError):
        M.ref_erode_horizontal_sweep(np.zeros((2, 2), np.int32), 0)
