"""CPU checks of the C ABI boundary: the library loads without a GPU and
exports every function include/split_hip.h declares; the C++ drop-in headers
compile against it."""
import re
import subprocess
from pathlib import Path

import pytest

import split_amd as sa

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "split_hip.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(split_hip_\w+)\s*\(", text)))


def test_header_declares_what_python_binds():
    assert declared_functions() == sorted(sa.CABI)


def test_library_exports_every_declared_symbol():
    lib = sa.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(sa.HIP_LIB)],
                         capture_output=True, text=True, check=True).stdout
    for name in declared_functions():
        assert re.search(rf"\bT {name}\b", out), name


def test_host_only_entry_points_without_gpu():
    lib = sa.lib()
    assert lib.split_hip_version().decode() == "0.1.0"
    assert lib.split_hip_status_string(2).decode() == "temp buffer too small"
    assert lib.split_hip_label_points_workspace(1000, 8) == 0
    # CC: linked-root bitmap (N/32 words + 2) and the 64x64 tiles' edge
    # arrays (4 x 64 ints per tile), each block 256-aligned, after a state slot
    n = 4096 * 4096
    ws = lib.split_hip_connected_components_workspace(4096, 4096)
    assert n // 8 + 4 * 4096 * 256 <= ws < n // 8 + 4 * 4096 * 256 + 4096
    ws = lib.split_hip_cluster_workspace(4096 * 4096, 32, 3)
    n = 4096 * 4096
    # per-pixel Lloyd state (4 B) + interleaved pixel copy (16 B) + small tables
    assert 20 * n <= ws < 20 * n + (1 << 20)
    assert lib.split_hip_compress_labels_workspace(1 << 20, 1 << 20) < 4 << 20
    # segment means (k beyond the register kernels): ~28 bytes per label
    ws = lib.split_hip_calculate_centroids_workspace(1 << 20, 200000)
    assert 28 * 200000 <= ws < 28 * 200000 + 4096
    # batch workspace: sizes only (buffers may be null), k limited to 64
    descs = (sa.ImageDesc * 3)()
    for i, (w, h, k) in enumerate([(2048, 2048, 64), (37, 53, 4), (1, 1, 1)]):
        descs[i] = sa.ImageDesc(None, None, None, None, w, h, k, 0)
    ws = lib.split_hip_segment_batch_workspace(descs, 3)
    n = 2048 * 2048 + 37 * 53 + 1
    # state 4 B/px + local-root, linked-root and prefix words (3 x 1/8 B/px)
    # + the CC tiles' edge columns (1/4 B/px) + the chunk table (512 KiB)
    assert 4 * n + 5 * n // 8 + (1 << 19) <= ws < 4 * n + 5 * n // 8 + (1 << 19) + (1 << 16)
    descs[1].k = 65
    assert lib.split_hip_segment_batch_workspace(descs, 3) == 0
    descs[1].k = 4
    descs[2].width = 0
    assert lib.split_hip_segment_batch_workspace(descs, 3) == 0


def test_argument_errors_need_no_gpu():
    lib = sa.lib()
    # null pointers and bad sizes are rejected before any HIP call
    assert lib.split_hip_label_points(None, 8, None, 100, None, None, 0, None) == 1
    assert "null" in lib.split_hip_last_error().decode()
    assert lib.split_hip_connected_components(None, 4, 4, None, 0, None, 10, 4, None) == 1


def test_cpp_dropin_headers_compile(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text("""
#include "split/device/kmeans/kmeans.cuh"
#include "split/device/ccl/connected_components.cuh"
#include "split/device/ccl/compress_labels.cuh"
#include "split/device/color/conversion.cuh"
#include "split/device/color/beta_feature.cuh"
#include "split/host/image_transfer.h"
#include "split/device/morph/erode.cuh"
#include "split/device/device_error.h"
using namespace split::device;
int main() {
  float* p = nullptr; float* c = nullptr; int* l = nullptr; int* s = nullptr;
  const int n = 64 * 64, k = 5;
  const_view2d<split::real> pts(3, n, p);
  view2d<split::real> cent(3, k, c);
  view1d<int> labels(l, n);
  device_ptr<void> temp(nullptr);
  (void)kmeans::cluster_workspace(n, k, 3);
  if (n < 0) {
    kmeans::uniform_random_initialize(pts, cent, 1u);
    auto r = kmeans::cluster(pts, cent, labels, temp, 100, 1e-3f);
    (void)r;
    ccl::connected_components(const_view2d<int>(64, 64, l), temp, view1d<int>(s, n), 100);
    (void)ccl::compress_labels(view1d<int>(s, n), temp);
    kmeans::calculate_centroids(const_view1d<int>(labels), pts, cent, temp);
    view2d<split::real> lab(3, n, p);
    color::convert_color_space(pts, lab, color::rgb_to_lab<>());
    color::convert_color_space(lab, lab, color::lab_to_rgb<>{});
    color::beta_feature(lab, lab.row(0));
    split::host::make_device_image(p, lab);
    split::host::make_host_image(lab, p);
    morph::erode(view2d<int>(64, 64, s), 10);
    morph::erode_horizontal_sweep(view2d<int>(64, 64, s));
  }
  return 0;
}
""")
    exe = tmp_path / "use"
    subprocess.run(["g++", "-std=c++17", f"-I{ROOT / 'include'}", str(src),
                    str(sa.HIP_LIB), "-o", str(exe), f"-Wl,-rpath,{sa.LIB_DIR}"],
                   check=True)
    subprocess.run([str(exe)], check=True)
