"""split_amd — MI355X (gfx950) segmentation hot path of nitronoid/split.

The compute path is the C-ABI library ``split_amd/_lib/libsplit_hip.so``
(hand-written HIP kernels); this module only loads it with ctypes and passes
device pointers from torch tensors (torch provides device memory and the
stream, nothing else).  There is no CPU fallback: if the library is missing,
every entry point raises ``SplitLibraryError``.

Entry points mirror the reference's C++ API (split/include/split/device/
kmeans/*.cuh, ccl/{connected_components,compress_labels}.cuh):

    kmeans.label_points(centroids, points, labels)
    kmeans.calculate_centroids(labels, points, centroids)
    kmeans.cluster(points, centroids, labels, max_iter, threshold) -> (iters, shift)
    kmeans.uniform_random_initialize(points, centroids, seed)
    ccl.connected_components(labels_hw, out, connectivity=4)
    ccl.compress_labels(labels) -> number of distinct labels
    segment(points, width, height, centroids, ...) -> SegmentResult
    segment_batch([{points, width, height, centroids, labels, segments}, ...])

Layouts: points (3, N) float32 SoA planes, centroids (3, K) float32,
labels int32 — all contiguous CUDA (HIP) tensors.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from pathlib import Path
from types import SimpleNamespace

LIB_DIR = Path(__file__).resolve().parent / "_lib"
HIP_LIB = LIB_DIR / "libsplit_hip.so"
GEN_LIB = LIB_DIR / "libsplit_gen.so"


class SplitLibraryError(RuntimeError):
    """The gfx950 library is missing or failed to load."""


class SplitError(RuntimeError):
    """A C-ABI call returned a nonzero status."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


class ClusterResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("shift", ctypes.c_float)]


class ImageDesc(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("centroids", ctypes.c_void_p),
                ("labels", ctypes.c_void_p), ("segments", ctypes.c_void_p),
                ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("k", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class GenParams(ctypes.Structure):
    """split_gen_params (include/split_gen.h)."""
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("sites", ctypes.c_int32),
        ("shading", ctypes.c_int32),
        ("octaves", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("base_cell", ctypes.c_double),
        ("height_amp", ctypes.c_double),
        ("noise_sigma", ctypes.c_double),
        ("noise_uniform", ctypes.c_double),
    ]


class GenDesc(ctypes.Structure):
    _fields_ = [("planes", ctypes.c_void_p), ("width", ctypes.c_int32),
                ("height", ctypes.c_int32), ("params", GenParams)]


class BatchResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("shift", ctypes.c_float),
                ("num_segments", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_c_void_p = ctypes.c_void_p
_c_size_t = ctypes.c_size_t
_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_u64 = ctypes.c_uint64
_c_float = ctypes.c_float

# name -> (restype, argtypes); exactly the declarations of include/split_hip.h
CABI = {
    "split_hip_last_error": (ctypes.c_char_p, []),
    "split_hip_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "split_hip_version": (ctypes.c_char_p, []),
    "split_hip_label_points_workspace": (_c_size_t, [_c_i64, _c_i32]),
    "split_hip_label_points": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p, _c_i64,
                                              _c_void_p, _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_label_image": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p, _c_i32, _c_i32,
                                             _c_void_p, _c_void_p]),
    "split_hip_calculate_centroids_workspace": (_c_size_t, [_c_i64, _c_i32]),
    "split_hip_calculate_centroids": (ctypes.c_int, [_c_void_p, _c_void_p, _c_i64, _c_void_p,
                                                     _c_i32, _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_calculate_centroids_noalloc": (ctypes.c_int, [_c_void_p, _c_void_p, _c_i64,
                                                             _c_void_p, _c_i32, _c_void_p]),
    "split_hip_propagate_centroids": (ctypes.c_int, [_c_void_p, _c_i64, _c_void_p, _c_i64,
                                                     _c_void_p, _c_void_p]),
    "split_hip_cluster_workspace": (_c_size_t, [_c_i64, _c_i32, _c_i32]),
    "split_hip_cluster": (ctypes.c_int, [_c_void_p, _c_i64, _c_void_p, _c_i32, _c_void_p,
                                         _c_void_p, _c_size_t, _c_i32, _c_float,
                                         ctypes.POINTER(ClusterResult), _c_void_p]),
    "split_hip_cluster_image": (ctypes.c_int, [_c_void_p, _c_i32, _c_i32, _c_void_p, _c_i32,
                                               _c_void_p, _c_void_p, _c_size_t, _c_i32,
                                               _c_float, ctypes.POINTER(ClusterResult),
                                               _c_void_p]),
    "split_hip_uniform_random_initialize": (ctypes.c_int, [_c_void_p, _c_i64, _c_void_p,
                                                           _c_i32, _c_u64, _c_void_p]),
    "split_hip_connected_components_workspace": (_c_size_t, [_c_i32, _c_i32]),
    "split_hip_connected_components": (ctypes.c_int, [_c_void_p, _c_i32, _c_i32, _c_void_p,
                                                      _c_size_t, _c_void_p, _c_i32, _c_i32,
                                                      _c_void_p]),
    "split_hip_compress_labels_workspace": (_c_size_t, [_c_i64, _c_i64]),
    "split_hip_compress_labels": (ctypes.c_int, [_c_void_p, _c_i64, _c_void_p, _c_size_t,
                                                 ctypes.POINTER(_c_i32), _c_void_p]),
    "split_hip_segment_workspace": (_c_size_t, [_c_i32, _c_i32, _c_i32]),
    "split_hip_segment": (ctypes.c_int, [_c_void_p, _c_i32, _c_i32, _c_void_p, _c_i32,
                                         _c_void_p, _c_void_p, _c_void_p, _c_size_t, _c_i32,
                                         _c_float, _c_i32, _c_i32,
                                         ctypes.POINTER(ClusterResult),
                                         ctypes.POINTER(_c_i32), _c_void_p]),
    "split_hip_segment_batch_workspace": (_c_size_t, [ctypes.POINTER(ImageDesc), _c_i32]),
    "split_hip_segment_batch": (ctypes.c_int, [ctypes.POINTER(ImageDesc), _c_i32, _c_void_p,
                                               _c_size_t, _c_i32, _c_float, _c_i32, _c_i32,
                                               _c_i32, ctypes.POINTER(BatchResult),
                                               _c_void_p]),
    "split_hip_ws_param_count": (_c_size_t, []),
    "split_hip_ws_workspace": (_c_size_t, [_c_i32]),
    "split_hip_ws_pack": (ctypes.c_int, [ctypes.POINTER(ImageDesc), _c_i32, _c_void_p, _c_void_p,
                                         _c_size_t, _c_void_p]),
    "split_hip_ws_coarse": (ctypes.c_int, [ctypes.POINTER(ImageDesc), _c_i32, _c_void_p,
                                           _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_ws_forward": (ctypes.c_int, [_c_void_p, _c_void_p, _c_void_p, _c_i32, _c_void_p,
                                            _c_size_t, _c_void_p, _c_void_p]),
    "split_hip_ws_backward": (ctypes.c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_i32,
                                             _c_i32, _c_void_p, _c_size_t, _c_void_p,
                                             _c_void_p, _c_void_p]),
    "split_hip_ws_adam": (ctypes.c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_i32,
                                         _c_float, _c_float, _c_float, _c_float, _c_void_p]),
    "split_hip_ws_predict": (ctypes.c_int, [_c_void_p, ctypes.POINTER(ImageDesc), _c_i32,
                                            _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_ws_targets": (ctypes.c_int, [_c_void_p, _c_i32, _c_void_p, _c_void_p]),
    "split_hip_convert_color_space": (ctypes.c_int, [_c_void_p, _c_void_p, _c_i64, _c_i32,
                                                     _c_void_p]),
    "split_hip_beta_feature_workspace": (_c_size_t, []),
    "split_hip_beta_feature": (ctypes.c_int, [_c_void_p, _c_i64, _c_void_p, _c_float,
                                              _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_lab_beta_feature": (ctypes.c_int, [_c_void_p, _c_i64, _c_void_p, _c_void_p,
                                                  _c_float, _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_interleaved_workspace": (_c_size_t, [_c_i64, _c_i32]),
    "split_hip_upload_interleaved": (ctypes.c_int, [_c_void_p, _c_i64, _c_i32, _c_void_p,
                                                    _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_download_interleaved": (ctypes.c_int, [_c_void_p, _c_i64, _c_i32, _c_void_p,
                                                      _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_deinterleave": (ctypes.c_int, [_c_void_p, _c_i64, _c_i32, _c_void_p,
                                              _c_void_p]),
    "split_hip_interleave": (ctypes.c_int, [_c_void_p, _c_i64, _c_i32, _c_void_p, _c_void_p]),
    "split_hip_quantize_interleaved": (ctypes.c_int, [_c_void_p, _c_i64, _c_i32, _c_void_p,
                                                      _c_void_p]),
    "split_hip_erode_workspace": (_c_size_t, [_c_i32, _c_i32]),
    "split_hip_erode_horizontal_sweep": (ctypes.c_int, [_c_void_p, _c_i32, _c_i32, _c_i32,
                                                        _c_void_p, _c_size_t, _c_void_p]),
    "split_hip_erode": (ctypes.c_int, [_c_void_p, _c_i32, _c_i32, _c_i32, _c_void_p,
                                       _c_size_t, _c_void_p]),
    "split_hip_inertia_workspace": (_c_size_t, []),
    "split_hip_inertia": (ctypes.c_int, [_c_void_p, _c_i64, _c_void_p, _c_i32, _c_void_p,
                                         _c_size_t, ctypes.POINTER(ctypes.c_double),
                                         _c_void_p]),
    "split_hip_gen_images_workspace": (_c_size_t, [ctypes.POINTER(GenDesc), _c_i32]),
    "split_hip_gen_images": (ctypes.c_int, [ctypes.POINTER(GenDesc), _c_i32, _c_void_p,
                                            _c_size_t, _c_void_p]),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Load libsplit_hip.so (once).  Raises SplitLibraryError if absent."""
    global _lib
    if _lib is None:
        if not HIP_LIB.exists():
            raise SplitLibraryError(
                f"{HIP_LIB} is missing: build it with `python -m split_amd.build` "
                "(there is no CPU fallback)")
        try:
            handle = ctypes.CDLL(str(HIP_LIB))
        except OSError as e:  # pragma: no cover - depends on the machine
            raise SplitLibraryError(f"cannot load {HIP_LIB}: {e}") from e
        for name, (res, args) in CABI.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def _check(status: int, what: str) -> None:
    if status != 0:
        L = lib()
        raise SplitError(status, f"{what}: {L.split_hip_status_string(status).decode()}: "
                                 f"{L.split_hip_last_error().decode()}")


def _torch():
    import torch
    return torch


def _stream() -> int:
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream


def _dev(t, dtype, what: str) -> int:
    torch = _torch()
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{what} must be a CUDA (HIP) tensor")
    if t.dtype != dtype:
        raise TypeError(f"{what} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    return t.data_ptr()


def _temp(nbytes: int, like):
    torch = _torch()
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=like.device)


def _points(points) -> tuple[int, int]:
    torch = _torch()
    if points.dim() != 2 or points.shape[0] != 3:
        raise ValueError("points must be (3, N)")
    return _dev(points, torch.float32, "points"), points.shape[1]


def _cent(centroids) -> tuple[int, int]:
    torch = _torch()
    if centroids.dim() != 2 or centroids.shape[0] != 3:
        raise ValueError("centroids must be (3, K)")
    return _dev(centroids, torch.float32, "centroids"), centroids.shape[1]


# --------------------------------------------------------------- k-means ---

def label_points(centroids, points, labels, temp=None, *, width: int = 0,
                 height: int = 0) -> None:
    """split::device::kmeans::label_points (label.cu:74-106).  With width and
    height (width*height == N) the kernel walks 16x16 tiles of the image."""
    torch = _torch()
    c, k = _cent(centroids)
    p, n = _points(points)
    lb = _dev(labels, torch.int32, "labels")
    if labels.numel() != n:
        raise ValueError("labels must hold N entries")
    if width:
        if width * height != n:
            raise ValueError("width*height must equal N")
        _check(lib().split_hip_label_image(c, k, p, width, height, lb, _stream()),
               "label_image")
    else:
        _check(lib().split_hip_label_points(c, k, p, n, lb, None, 0, _stream()),
               "label_points")


def calculate_centroids(labels, points, centroids, temp=None) -> None:
    """split::device::kmeans::calculate_centroids (centroids.cu:22-101)."""
    torch = _torch()
    p, n = _points(points)
    c, k = _cent(centroids)
    lb = _dev(labels, torch.int32, "labels")
    if labels.numel() != n:
        raise ValueError("labels must hold N entries")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_calculate_centroids_workspace(n, k), points)
    _check(L.split_hip_calculate_centroids(lb, p, n, c, k, temp.data_ptr(), temp.numel(),
                                           _stream()), "calculate_centroids")


def calculate_centroids_noalloc(labels, points, centroids) -> None:
    """calculate_centroids with no temp and no device allocation: the kernels
    use the library's static per-GPU scratch (calls that overlap on one GPU
    race)."""
    torch = _torch()
    p, n = _points(points)
    c, k = _cent(centroids)
    lb = _dev(labels, torch.int32, "labels")
    if labels.numel() != n:
        raise ValueError("labels must hold N entries")
    _check(lib().split_hip_calculate_centroids_noalloc(lb, p, n, c, k, _stream()),
           "calculate_centroids_noalloc")


def propagate_centroids(labels, centroids, points) -> None:
    """split::device::kmeans::propagate_centroids (propagation.cu:9-24):
    points[:, i] = centroids[:, labels[i]] (labels outside [0, K) untouched)."""
    torch = _torch()
    p, n = _points(points)
    c, k = _cent(centroids)
    lb = _dev(labels, torch.int32, "labels")
    if labels.numel() != n:
        raise ValueError("labels must hold N entries")
    _check(lib().split_hip_propagate_centroids(lb, n, c, k, p, _stream()),
           "propagate_centroids")


def cluster(points, centroids, labels, max_iter: int, threshold: float = 1.0,
            temp=None, *, width: int = 0, height: int = 0) -> tuple[int, float]:
    """split::device::kmeans::cluster (cluster.cu:27-93).  Returns
    (iterations, last shift); host-synchronous like the reference.  With
    width and height the kernel walks 16x16 tiles of the image."""
    torch = _torch()
    p, n = _points(points)
    c, k = _cent(centroids)
    lb = _dev(labels, torch.int32, "labels")
    if labels.numel() != n:
        raise ValueError("labels must hold N entries")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_cluster_workspace(n, k, 3), points)
    r = ClusterResult()
    if width:
        if width * height != n:
            raise ValueError("width*height must equal N")
        _check(L.split_hip_cluster_image(p, width, height, c, k, lb, temp.data_ptr(),
                                         temp.numel(), int(max_iter), float(threshold),
                                         ctypes.byref(r), _stream()), "cluster_image")
    else:
        _check(L.split_hip_cluster(p, n, c, k, lb, temp.data_ptr(), temp.numel(),
                                   int(max_iter), float(threshold), ctypes.byref(r),
                                   _stream()), "cluster")
    return int(r.iterations), float(r.shift)


def uniform_random_initialize(points, centroids, seed: int) -> None:
    """Seeded form of split::device::kmeans::uniform_random_initialize."""
    p, n = _points(points)
    c, k = _cent(centroids)
    _check(lib().split_hip_uniform_random_initialize(p, n, c, k, int(seed), _stream()),
           "uniform_random_initialize")


kmeans = SimpleNamespace(label_points=label_points, calculate_centroids=calculate_centroids,
                         calculate_centroids_noalloc=calculate_centroids_noalloc,
                         cluster=cluster, uniform_random_initialize=uniform_random_initialize,
                         propagate_centroids=propagate_centroids)


# ------------------------------------------------------------------- ccl ---

def connected_components(labels_hw, out, connectivity: int = 4, max_iterations: int = 10,
                         temp=None) -> None:
    """split::device::ccl::connected_components (connected_components.cu:80-181):
    out[p] = largest linear index of p's component."""
    torch = _torch()
    if labels_hw.dim() != 2:
        raise ValueError("labels must be (H, W)")
    h, w = labels_hw.shape
    li = _dev(labels_hw, torch.int32, "labels")
    oi = _dev(out, torch.int32, "out")
    if out.numel() != h * w:
        raise ValueError("out must hold H*W entries")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_connected_components_workspace(w, h), out)
    _check(L.split_hip_connected_components(li, w, h, temp.data_ptr(), temp.numel(), oi,
                                            int(max_iterations), int(connectivity),
                                            _stream()),
           "connected_components")


def compress_labels(labels, temp=None) -> int:
    """split::device::ccl::compress_labels (compress_labels.cu:10-35).  In
    place; returns the number of distinct values (host-synchronous)."""
    torch = _torch()
    li = _dev(labels, torch.int32, "labels")
    n = labels.numel()
    L = lib()
    if temp is None:
        # the reference's contract: a temp of N ints, whose bitmap holds a
        # value span of 32 N (any connected-components output); a wider span
        # returns SPLIT_HIP_ERR_RANGE, and a caller passes a larger temp
        temp = _temp(max(L.split_hip_compress_labels_workspace(n, n), 4 * n), labels)
    cnt = _c_i32(0)
    _check(L.split_hip_compress_labels(li, n, temp.data_ptr(), temp.numel(),
                                       ctypes.byref(cnt), _stream()), "compress_labels")
    return int(cnt.value)


ccl = SimpleNamespace(connected_components=connected_components,
                      compress_labels=compress_labels)


# ----------------------------------------------------------------- morph ---

def _erode_call(fn: str, labels_hw, niterations: int, temp):
    torch = _torch()
    if labels_hw.dim() != 2:
        raise ValueError("labels must be (H, W)")
    h, w = labels_hw.shape
    lb = _dev(labels_hw, torch.int32, "labels")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_erode_workspace(w, h), labels_hw)
    _check(getattr(L, fn)(lb, w, h, int(niterations), temp.data_ptr(), temp.numel(),
                          _stream()), fn[len("split_hip_"):])


def erode(labels_hw, niterations: int = 1, temp=None) -> None:
    """split::device::morph::erode (erode.cu:141-159), in place: a pixel
    keeps its label iff the (2n+1)^2 square around it lies in the image and
    carries one label; other pixels become -1."""
    _erode_call("split_hip_erode", labels_hw, niterations, temp)


def erode_horizontal_sweep(labels_hw, niterations: int = 1, temp=None) -> None:
    """split::device::morph::erode_horizontal_sweep (erode.cu:57-139), in
    place: the row sweep alone."""
    _erode_call("split_hip_erode_horizontal_sweep", labels_hw, niterations, temp)


morph = SimpleNamespace(erode=erode, erode_horizontal_sweep=erode_horizontal_sweep)


# ----------------------------------------------------------------- color ---

# split_hip_color_conversion values, named after the reference's functors
# (split/include/split/device/color/conversion.cuh:15-62)
CONVERSIONS = {"rgb_to_xyz": 0, "xyz_to_rgb": 1, "xyz_to_lab": 2, "lab_to_xyz": 3,
               "rgb_to_lab": 4, "lab_to_rgb": 5, "rgb_to_ic": 6}


def convert_color_space(points_in, points_out, converter: str) -> None:
    """split::device::color::convert_color_space (conversion.cuh:67-82) with
    the converter named as the reference's functor ("rgb_to_lab", ...).
    points_out may be points_in (in place)."""
    if converter not in CONVERSIONS:
        raise ValueError(f"unknown converter {converter!r}; one of {sorted(CONVERSIONS)}")
    pi, n = _points(points_in)
    po, n2 = _points(points_out)
    if n2 != n:
        raise ValueError("in and out must both be (3, N)")
    _check(lib().split_hip_convert_color_space(pi, po, n, CONVERSIONS[converter], _stream()),
           "convert_color_space")


def beta_feature(lab_points, beta, mu: float = 1e5, temp=None) -> None:
    """split::device::color::beta_feature (beta_feature.cu:48-78): beta (N
    floats, may be a row of lab_points) from the (3, N) Lab points."""
    torch = _torch()
    p, n = _points(lab_points)
    b = _dev(beta, torch.float32, "beta")
    if beta.numel() != n:
        raise ValueError("beta must hold N entries")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_beta_feature_workspace(), lab_points)
    _check(L.split_hip_beta_feature(p, n, b, float(mu), temp.data_ptr(), temp.numel(),
                                    _stream()), "beta_feature")


def lab_beta_feature(rgb_points, features, luminance=None, mu: float = 1e5,
                     temp=None) -> None:
    """The sample's feature sequence (main.cu:216-225): Lab of rgb_points into
    features (3, N), L copied into luminance (N, optional), beta over L."""
    torch = _torch()
    p, n = _points(rgb_points)
    f, n2 = _points(features)
    if n2 != n:
        raise ValueError("rgb_points and features must both be (3, N)")
    lum = None
    if luminance is not None:
        lum = _dev(luminance, torch.float32, "luminance")
        if luminance.numel() != n:
            raise ValueError("luminance must hold N entries")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_beta_feature_workspace(), rgb_points)
    _check(L.split_hip_lab_beta_feature(p, n, f, lum, float(mu), temp.data_ptr(),
                                        temp.numel(), _stream()), "lab_beta_feature")


color = SimpleNamespace(convert_color_space=convert_color_space, beta_feature=beta_feature,
                        lab_beta_feature=lab_beta_feature, CONVERSIONS=CONVERSIONS)


# -------------------------------------------------------------------- io ---

def _hwc_host(host):
    """(ptr, n, channels) of a C-contiguous float32 numpy array or CPU tensor
    of shape (..., C): the last axis is the channel."""
    import numpy as np
    torch = _torch()
    if isinstance(host, torch.Tensor):
        if host.is_cuda or host.dtype != torch.float32 or not host.is_contiguous():
            raise TypeError("host image must be a contiguous float32 CPU tensor")
        c = host.shape[-1]
        return host.data_ptr(), host.numel() // c, c
    if not isinstance(host, np.ndarray) or host.dtype != np.float32 or \
            not host.flags["C_CONTIGUOUS"]:
        raise TypeError("host image must be a C-contiguous float32 array")
    c = host.shape[-1]
    return host.ctypes.data, host.size // c, c


def make_device_image(host_hwc, planes, temp=None) -> None:
    """make_device_image (samples/src/main.cu:98-114): interleaved host image
    (..., C) -> device planes (C, N).  Host-synchronous."""
    torch = _torch()
    h, n, c = _hwc_host(host_hwc)
    d = _dev(planes, torch.float32, "planes")
    if planes.dim() != 2 or tuple(planes.shape) != (c, n):
        raise ValueError(f"planes must be ({c}, {n})")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_interleaved_workspace(n, c), planes)
    _check(L.split_hip_upload_interleaved(h, n, c, d, temp.data_ptr(), temp.numel(),
                                          _stream()), "upload_interleaved")


def make_host_image(planes, host_hwc, temp=None) -> None:
    """make_host_image (samples/src/main.cu:116-132): device planes (C, N) ->
    interleaved host image (..., C).  Host-synchronous."""
    torch = _torch()
    h, n, c = _hwc_host(host_hwc)
    d = _dev(planes, torch.float32, "planes")
    if planes.dim() != 2 or tuple(planes.shape) != (c, n):
        raise ValueError(f"planes must be ({c}, {n})")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_interleaved_workspace(n, c), planes)
    _check(L.split_hip_download_interleaved(d, n, c, h, temp.data_ptr(), temp.numel(),
                                            _stream()), "download_interleaved")


def deinterleave(hwc, planes) -> None:
    """Device (N, C) or (..., C) interleaved -> (C, N) planes."""
    torch = _torch()
    c = planes.shape[0]
    n = planes.shape[1]
    if hwc.numel() != n * c or hwc.shape[-1] != c:
        raise ValueError("hwc must hold N pixels of C channels")
    _check(lib().split_hip_deinterleave(_dev(hwc, torch.float32, "hwc"), n, c,
                                        _dev(planes, torch.float32, "planes"), _stream()),
           "deinterleave")


def interleave(planes, hwc) -> None:
    """Device (C, N) planes -> (..., C) interleaved."""
    torch = _torch()
    c, n = planes.shape
    if hwc.numel() != n * c or hwc.shape[-1] != c:
        raise ValueError("hwc must hold N pixels of C channels")
    _check(lib().split_hip_interleave(_dev(planes, torch.float32, "planes"), n, c,
                                      _dev(hwc, torch.float32, "hwc"), _stream()),
           "interleave")


def quantize(planes, out) -> None:
    """stbi::quantize (stbi_raii.cpp:135-143) of device planes (C, N) into a
    device uint8 (..., C) interleaved image."""
    torch = _torch()
    c, n = planes.shape
    if out.numel() != n * c or out.shape[-1] != c:
        raise ValueError("out must hold N pixels of C channels")
    _check(lib().split_hip_quantize_interleaved(_dev(planes, torch.float32, "planes"), n, c,
                                                _dev(out, torch.uint8, "out"), _stream()),
           "quantize_interleaved")


io = SimpleNamespace(make_device_image=make_device_image, make_host_image=make_host_image,
                     deinterleave=deinterleave, interleave=interleave, quantize=quantize)


# --------------------------------------------------------------- segment ---

@dataclass
class SegmentResult:
    iterations: int
    shift: float
    num_segments: int


def segment_workspace(width: int, height: int, k: int) -> int:
    return int(lib().split_hip_segment_workspace(width, height, k))


def segment(points, width: int, height: int, centroids, labels, segments, *,
            max_iter: int = 100, threshold: float = 1e-4, connectivity: int = 8,
            compress: bool = True, temp=None) -> SegmentResult:
    """cluster -> connected_components -> compress_labels, stream-ordered
    (the caller sequence of samples/src/main.cu:240-252)."""
    torch = _torch()
    p, n = _points(points)
    if n != width * height:
        raise ValueError("points must hold width*height pixels")
    c, k = _cent(centroids)
    lb = _dev(labels, torch.int32, "labels")
    sg = _dev(segments, torch.int32, "segments")
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_segment_workspace(width, height, k), points)
    r = ClusterResult()
    ns = _c_i32(0)
    _check(L.split_hip_segment(p, width, height, c, k, lb, sg, temp.data_ptr(), temp.numel(),
                               int(max_iter), float(threshold), int(connectivity),
                               int(bool(compress)), ctypes.byref(r), ctypes.byref(ns),
                               _stream()), "segment")
    return SegmentResult(int(r.iterations), float(r.shift), int(ns.value))


def _batch_descs(images):
    torch = _torch()
    descs = (ImageDesc * len(images))()
    for i, im in enumerate(images):
        p, n = _points(im["points"])
        w, h = int(im["width"]), int(im["height"])
        if n != w * h:
            raise ValueError(f"image {i}: points must hold width*height pixels")
        c, k = _cent(im["centroids"])
        lb = _dev(im["labels"], torch.int32, "labels")
        seg = im.get("segments")
        sg = None if seg is None else _dev(seg, torch.int32, "segments")
        descs[i] = ImageDesc(p, c, lb, sg, w, h, k, 0)
    return descs


def segment_batch_workspace(images) -> int:
    """Bytes of temp for segment_batch over `images` (dicts as below)."""
    descs = _batch_descs(images)
    return int(lib().split_hip_segment_batch_workspace(descs, len(images)))


def segment_batch(images, *, max_iter: int = 100, threshold: float = 1e-4,
                  connectivity: int = 8, compress: bool = True, use_graph: bool = True,
                  temp=None) -> list[SegmentResult]:
    """segment() of every image in one batched, graph-replayed launch sequence.

    images: sequence of dicts with keys points (3, w*h), width, height,
    centroids (3, k) in/out, labels (w*h,), segments (w*h,; may be absent
    with connectivity 0, the clustering alone); k <= 64.
    Each image's outputs equal segment() of that image alone."""
    descs = _batch_descs(images)
    L = lib()
    count = len(images)
    if temp is None:
        temp = _temp(L.split_hip_segment_batch_workspace(descs, count),
                     images[0]["points"])
    res = (BatchResult * count)()
    _check(L.split_hip_segment_batch(descs, count, temp.data_ptr(), temp.numel(),
                                     int(max_iter), float(threshold), int(connectivity),
                                     int(bool(compress)), int(bool(use_graph)), res,
                                     _stream()), "segment_batch")
    return [SegmentResult(int(r.iterations), float(r.shift), int(r.num_segments))
            for r in res]


def gen_images(specs, outs, temp=None) -> None:
    """Seeded synthetic images on the device: outs[i] ((3, w*h) float32 CUDA
    tensor) <- the image of specs[i] (a workloads.ImageSpec)."""
    count = len(specs)
    descs = (GenDesc * count)()
    for i, (sp, o) in enumerate(zip(specs, outs)):
        p, n = _points(o)
        if n != sp.width * sp.height:
            raise ValueError(f"image {i}: output must hold width*height pixels")
        descs[i] = GenDesc(p, sp.width, sp.height, sp.params())
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_gen_images_workspace(descs, count), outs[0])
    _check(L.split_hip_gen_images(descs, count, temp.data_ptr(), temp.numel(), _stream()),
           "gen_images")


def inertia(points, centroids, temp=None) -> float:
    """sum_i min_j |p_i - c_j|^2 (float distances, double sum)."""
    p, n = _points(points)
    c, k = _cent(centroids)
    L = lib()
    if temp is None:
        temp = _temp(L.split_hip_inertia_workspace(), points)
    out = ctypes.c_double(0.0)
    _check(L.split_hip_inertia(p, n, c, k, temp.data_ptr(), temp.numel(), ctypes.byref(out),
                               _stream()), "inertia")
    return float(out.value)


def version() -> str:
    return lib().split_hip_version().decode()
