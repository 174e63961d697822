"""ctypes binding of the CPU oracle (oracle/oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker or the CPU baseline.
The product path (split_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
# the same code built with OpenMP: only the CPU baseline's all-cores timing
# loads it (set OMP_NUM_THREADS before the first call)
LIB_OMP = HERE / "_build" / "liboracle_omp.so"

_lib = None
_lib_omp = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_F = ctypes.c_float

_SIGS = {
    "oracle_label_f64": (None, [_P, _I64, _P, _I, _P, _P]),
    "oracle_centroids_f64": (None, [_P, _P, _I64, _I, _P, _P]),
    "oracle_lloyd_f64": (None, [_P, _I64, _P, _I, _I, _P, _P]),
    "oracle_inertia": (ctypes.c_double, [_P, _I64, _P, _I, _P]),
    "oracle_cc_uf": (_I64, [_P, _I, _I, _I, _P]),
    "oracle_compress": (_I64, [_P, _I64]),
    "oracle_sample_indices": (None, [ctypes.c_uint64, _I64, _I, _P]),
    "oracle_threads": (_I, []),
    "ref_label_points": (None, [_P, _I, _P, _I64, _P]),
    "ref_calculate_centroids": (None, [_P, _P, _I64, _P, _I]),
    "ref_cluster": (_I, [_P, _I64, _P, _I, _P, _I, _F, ctypes.POINTER(_F)]),
    "ref_connected_components": (_I, [_P, _I, _I, _P, _I]),
    "ref_compress_labels": (_I64, [_P, _I64]),
}


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _load(path: Path) -> ctypes.CDLL:
    if not path.exists():
        build()
    h = ctypes.CDLL(str(path))
    for name, (res, args) in _SIGS.items():
        f = getattr(h, name)
        f.restype = res
        f.argtypes = args
    return h


_lib_ref = None

# The reference's host flags (split's common.pri:27-32): the CPU baseline's
# all-cores timing build.  -march=native must match the CPU that runs it, so
# that library is compiled where it is used, into a temporary directory, and
# never shipped.
REFERENCE_FLAGS = ["-Ofast", "-march=native", "-fopenmp", "-mfma", "-mavx2"]


def lib_reference_flags() -> ctypes.CDLL:
    """oracle.c built with the reference's host flags and OpenMP, for timing
    only (fast-math: not the checker)."""
    global _lib_ref
    if _lib_ref is None:
        import tempfile
        d = Path(tempfile.mkdtemp(prefix="oracle_ref_flags_"))
        out = d / "liboracle_refflags.so"
        subprocess.run(["gcc", *REFERENCE_FLAGS, "-fPIC", "-shared", "-std=c11",
                        str(HERE / "oracle.c"), "-o", str(out), "-lm"], check=True)
        _lib_ref = _load(out)
    return _lib_ref


def lib(parallel: bool = False) -> ctypes.CDLL:
    global _lib, _lib_omp
    if parallel:
        if _lib_omp is None:
            _lib_omp = _load(LIB_OMP)
        return _lib_omp
    if _lib is None:
        _lib = _load(LIB)
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def label_f64(pts, cent):
    pts, cent = _f32(pts), _f32(cent)
    n, k = pts.shape[1], cent.shape[1]
    labels = np.empty(n, np.int32)
    gap = np.empty(n, np.float64)
    lib().oracle_label_f64(pts.ctypes.data, n, cent.ctypes.data, k, labels.ctypes.data,
                           gap.ctypes.data)
    return labels, gap


def centroids_f64(labels, pts, k):
    pts, labels = _f32(pts), _i32(labels)
    out = np.empty(3 * k, np.float64)
    counts = np.empty(k, np.int64)
    lib().oracle_centroids_f64(labels.ctypes.data, pts.ctypes.data, pts.shape[1], k,
                               out.ctypes.data, counts.ctypes.data)
    return out.reshape(3, k), counts


def lloyd_f64(pts, init, iters):
    pts, init = _f32(pts), _f32(init)
    k = init.shape[1]
    cent = np.empty(3 * k, np.float64)
    labels = np.empty(pts.shape[1], np.int32)
    lib().oracle_lloyd_f64(pts.ctypes.data, pts.shape[1], init.ctypes.data, k, int(iters),
                           cent.ctypes.data, labels.ctypes.data)
    return cent.reshape(3, k), labels


def inertia(pts, cent, labels):
    pts, labels = _f32(pts), _i32(labels)
    cent = np.ascontiguousarray(cent, dtype=np.float32)
    return lib().oracle_inertia(pts.ctypes.data, pts.shape[1], cent.ctypes.data,
                                cent.shape[1], labels.ctypes.data)


def cc_uf(labels_hw, conn, handle=None):
    labels_hw = _i32(labels_hw)
    h, w = labels_hw.shape
    out = np.empty(h * w, np.int32)
    cnt = (handle or lib()).oracle_cc_uf(labels_hw.ctypes.data, w, h, int(conn), out.ctypes.data)
    return out, int(cnt)


def compress(labels, handle=None):
    a = _i32(labels).copy()
    cnt = (handle or lib()).oracle_compress(a.ctypes.data, a.size)
    return a, int(cnt)


def sample_indices(seed, n, k):
    idx = np.empty(k, np.int64)
    lib().oracle_sample_indices(int(seed) & (2**64 - 1), int(n), int(k), idx.ctypes.data)
    return idx


def ref_label_points(cent, pts):
    pts, cent = _f32(pts), _f32(cent)
    labels = np.empty(pts.shape[1], np.int32)
    lib().ref_label_points(cent.ctypes.data, cent.shape[1], pts.ctypes.data, pts.shape[1],
                           labels.ctypes.data)
    return labels


def ref_calculate_centroids(labels, pts, k):
    pts, labels = _f32(pts), _i32(labels)
    cent = np.empty(3 * k, np.float32)
    lib().ref_calculate_centroids(labels.ctypes.data, pts.ctypes.data, pts.shape[1],
                                  cent.ctypes.data, k)
    return cent.reshape(3, k)


def ref_cluster(pts, init, max_iter, threshold, parallel=False, handle=None):
    """handle: another build of the same code (the timing builds)."""
    pts = _f32(pts)
    cent = _f32(init).copy()
    labels = np.empty(pts.shape[1], np.int32)
    d = ctypes.c_float(0.0)
    it = (handle or lib(parallel)).ref_cluster(pts.ctypes.data, pts.shape[1], cent.ctypes.data,
                                   cent.shape[1], labels.ctypes.data, int(max_iter),
                                   float(threshold), ctypes.byref(d))
    return cent, labels, int(it), float(d.value)


def ref_connected_components(labels_hw, max_iter):
    labels_hw = _i32(labels_hw)
    h, w = labels_hw.shape
    out = np.empty(h * w, np.int32)
    it = lib().ref_connected_components(labels_hw.ctypes.data, w, h, out.ctypes.data,
                                        int(max_iter))
    return out, int(it)


def canonical_min(rep):
    """Relabel a component labelling so each component carries its smallest
    linear index (the canonical form both sides are compared in)."""
    rep = np.asarray(rep).ravel()
    n = rep.size
    idx = np.arange(n, dtype=np.int64)
    mins = np.full(int(rep.max()) + 1 if n else 1, np.iinfo(np.int64).max, dtype=np.int64)
    np.minimum.at(mins, rep, idx)
    return mins[rep].astype(np.int64)


def threads(parallel=False, handle=None):
    """Threads the (OpenMP) build runs its per-pixel loops on."""
    return int((handle or lib(parallel)).oracle_threads())
