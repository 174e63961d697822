"""CPU oracle of label erosion (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and bench.py's CPU leg) as the
checker of the gfx950 erosion kernels.  The product path never imports it.

ref_erode_horizontal_sweep restates split/src/split/device/morph/erode.cu:57-139
step by step: row-wise segments of equal labels from an adjacent difference
and a scan (:34-55, :67-73), segment sizes by reduce_by_key (:77-89), an
exclusive scan of the sizes (:90-92), sizes shrunk by 2n and clamped at 0
(:94-99), empty segments removed (:101-108), 1s scattered at base = start + n
and end = base + new size (:113-130), an inclusive scan (:133-135) and every
pixel whose scan value is even set to -1 (:136-138).  ref_erode restates
erode.cu:141-159 (sweep of the transposed labels, transposed back, then a
sweep of the rows).  For n = 0 the reference scatters to index N (past its
buffer) and its scan then drops every other run, so the restatement only
covers n >= 1.

erode_window is the independent definition the restatement is checked
against: a pixel survives a row sweep iff the 2n + 1 pixels centred on it lie
inside its row and carry its label.
"""
from __future__ import annotations

import numpy as np


def ref_erode_horizontal_sweep(labels, n):
    """erode.cu:57-139 on a (H, W) int array; returns a new array."""
    if n < 1:
        raise ValueError("the reference's sweep is defined for n >= 1 only")
    lab = np.asarray(labels)
    h, w = lab.shape
    flat = lab.ravel().astype(np.int64)
    npx = flat.size
    rows = np.arange(npx) // w
    key = np.ones(npx, dtype=np.int64)
    key[1:] = ((rows[1:] != rows[:-1]) | (flat[1:] != flat[:-1])).astype(np.int64)
    seg = np.cumsum(key) - 1
    sizes = np.bincount(seg)
    cum = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    new = np.maximum(sizes - 2 * n, 0)
    keep = new != 0
    marks = np.zeros(npx, dtype=np.int64)
    base = cum[keep] + n
    marks[base] = 1
    marks[base + new[keep]] = 1
    scan = np.cumsum(marks)
    out = np.where(scan % 2 == 0, -1, flat)
    return out.reshape(h, w).astype(np.int32)


def ref_erode(labels, n):
    """erode.cu:141-159: the sweep on the transposed labels, transposed
    back, then the sweep on the rows."""
    t = np.ascontiguousarray(np.asarray(labels).T)
    t = ref_erode_horizontal_sweep(t, n)
    return ref_erode_horizontal_sweep(np.ascontiguousarray(t.T), n)


def erode_window(labels, n, axis=1):
    """Independent definition of one sweep along `axis` (1: rows, 0:
    columns): keep a pixel iff its 2n + 1 window along the axis is inside
    the image and uniform."""
    lab = np.asarray(labels)
    if axis == 0:
        return np.ascontiguousarray(erode_window(lab.T, n, 1).T)
    h, w = lab.shape
    out = np.full((h, w), -1, dtype=np.int32)
    for c in range(n, w - n):
        win = lab[:, c - n:c + n + 1]
        same = np.all(win == lab[:, c:c + 1], axis=1)
        out[same, c] = lab[same, c]
    return out
