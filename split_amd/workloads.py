"""Seeded synthetic workloads of the segmentation path.

The images come from the host C generator (split_amd/csrc/gen.c, built into
split_amd/_lib/libsplit_gen.so): Voronoi colour regions, optionally
multiplied by Lambertian shading of a value-noise height map, plus noise,
clamped to [0, 1].  Identical parameters give bit-identical images on every
machine that runs the generator.

Workloads (name, size, clustering, connected components):
  plain_512   512x512, 64 sites, noise sigma 0.02, k=8, exactly 20 Lloyd
              iterations, 4-connectivity
  shaded_4096 4096x4096, 1000 sites x Lambertian shading (8 octaves),
              noise sigma 0.01, k=32, stop at shift < 1e-4 or 100
              iterations, 8-connectivity
  stress_8192 8192x8192, 20000 sites + uniform noise in [-0.05, 0.05],
              k=256; plus the 8192x8192 spiral/checkerboard label image for
              connected components only, 8-connectivity
  batch_256   256 images generated as shaded_4096 (shading, sigma 0.01),
              each side drawn from 128..2048, 50..2000 sites, k drawn from
              {4, 8, 16, 32, 64}; stop at shift < 1e-4 or 100 iterations,
              8-connectivity (see batch_workloads)
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace
from pathlib import Path

import numpy as np

from . import GEN_LIB, GenParams, SplitLibraryError


_gen = None


def gen_lib() -> ctypes.CDLL:
    global _gen
    if _gen is None:
        if not GEN_LIB.exists():
            raise SplitLibraryError(f"{GEN_LIB} is missing: run `python -m split_amd.build`")
        g = ctypes.CDLL(str(GEN_LIB))
        g.split_gen_image.restype = ctypes.c_int
        g.split_gen_image.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(GenParams)]
        g.split_gen_cc_stress.restype = ctypes.c_int
        g.split_gen_cc_stress.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        _gen = g
    return _gen


@dataclass(frozen=True)
class ImageSpec:
    width: int
    height: int
    seed: int
    sites: int
    shading: bool = False
    octaves: int = 8
    base_cell: float = 0.0      # 0 -> max(width, height) / 8
    height_amp: float = 0.0     # 0 -> base_cell / 8
    noise_sigma: float = 0.0
    noise_uniform: float = 0.0

    def params(self) -> GenParams:
        cell = self.base_cell or max(self.width, self.height) / 8.0
        amp = self.height_amp or cell / 8.0
        return GenParams(self.seed, self.sites, int(self.shading), self.octaves, 0, cell, amp,
                         self.noise_sigma, self.noise_uniform)


def make_image(spec: ImageSpec) -> np.ndarray:
    """(3, H*W) float32 SoA planes."""
    out = np.empty((3, spec.width * spec.height), dtype=np.float32)
    p = spec.params()
    rc = gen_lib().split_gen_image(out.ctypes.data, spec.width, spec.height, ctypes.byref(p))
    if rc != 0:
        raise RuntimeError("split_gen_image failed")
    return out


def make_cc_stress(width: int, height: int) -> np.ndarray:
    """(H, W) int32 spiral + checkerboard label image."""
    out = np.empty((height, width), dtype=np.int32)
    if gen_lib().split_gen_cc_stress(out.ctypes.data, width, height) != 0:
        raise RuntimeError("split_gen_cc_stress failed")
    return out


@dataclass(frozen=True)
class Workload:
    name: str
    image: ImageSpec
    k: int
    init_seed: int
    max_iter: int
    threshold: float      # < 0: run exactly max_iter iterations
    connectivity: int

    @property
    def n(self) -> int:
        return self.image.width * self.image.height


PLAIN_512 = Workload(
    "plain_512",
    ImageSpec(512, 512, seed=0x5A17_0001, sites=64, noise_sigma=0.02),
    k=8, init_seed=0x1417_0001, max_iter=20, threshold=-1.0, connectivity=4)

SHADED_4096 = Workload(
    "shaded_4096",
    ImageSpec(4096, 4096, seed=0x5A17_0002, sites=1000, shading=True, noise_sigma=0.01),
    k=32, init_seed=0x1417_0002, max_iter=100, threshold=1e-4, connectivity=8)

STRESS_8192 = Workload(
    "stress_8192",
    ImageSpec(8192, 8192, seed=0x5A17_0005, sites=20000, noise_uniform=0.05),
    k=256, init_seed=0x1417_0005, max_iter=100, threshold=1e-4, connectivity=8)

WORKLOADS = {w.name: w for w in (PLAIN_512, SHADED_4096, STRESS_8192)}


def scaled(w: Workload, width: int, height: int) -> Workload:
    """Same generator settings at another size (sites scaled by area)."""
    sites = max(1, round(w.image.sites * (width * height) / (w.image.width * w.image.height)))
    return replace(w, name=f"{w.name}@{width}x{height}",
                   image=replace(w.image, width=width, height=height, sites=sites,
                                 base_cell=w.image.base_cell or max(w.image.width,
                                                                    w.image.height) / 8.0))


BATCH_KS = (4, 8, 16, 32, 64)


def batch_workloads(count: int = 256, seed: int = 0x5A17_0003,
                    side_range: tuple[int, int] = (128, 2048)) -> list[Workload]:
    """The images of a batched run: sizes, site counts, k and seeds drawn
    from a seeded PCG64 stream, each image generated as shaded_4096."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = side_range
    out = []
    for i in range(count):
        w = int(rng.integers(lo, hi + 1))
        h = int(rng.integers(lo, hi + 1))
        sites = int(rng.integers(50, 2001))
        k = int(BATCH_KS[int(rng.integers(0, len(BATCH_KS)))])
        img_seed = int(rng.integers(0, 2**63))
        init_seed = int(rng.integers(0, 2**63))
        out.append(Workload(
            f"batch[{i}]",
            ImageSpec(w, h, seed=img_seed, sites=sites, shading=True, noise_sigma=0.01),
            k=k, init_seed=init_seed, max_iter=100, threshold=1e-4, connectivity=8))
    return out
