"""Build recipe for the gfx950 library, the C++ drop-in wrappers and the
synthetic-input generator.

    python -m split_amd.build            # everything, in-tree
    python -m split_amd.build --jobs 8

Outputs (git-ignored, shipped with the tree):
    split_amd/_lib/libsplit_hip.so   HIP kernels + C ABI (include/split_hip.h)
                                     + C++ drop-in wrappers (include/split/...)
    split_amd/_lib/libsplit_gen.so   seeded image generators (host C)
    split_amd/_lib/split_dropin_sequence
                                     the reference sample's call sequence
                                     against the drop-in headers
                                     (tests/dropin/main_sequence.cpp), run
                                     by the GPU tests
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "split_amd"
CSRC = PKG / "csrc"
LIB = PKG / "_lib"
OBJ = LIB / "obj"
INCLUDE = ROOT / "include"

ARCH = "gfx950"
HIP_SOURCES = ["kmeans.hip", "lloyd_tiles.hip", "ccl.hip", "batch.hip", "gen_device.hip", "warmstart.hip",
               "segmeans.hip", "color.hip", "io.hip", "erode.hip", "capi.hip"]
CXX_SOURCES = ["split_api.cpp"]
GEN_SOURCES = ["gen.c"]

# Packed FP32 VALU forms are switched off for all device code.  Inputs are
# finite by contract, so kernels run with the IEEE mode bit off and without
# NaN semantics: v_min/v_max need no canonicalising moves and fold into DPP.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-honor-nans", "-mno-amdgpu-ieee",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the gfx950 library cannot be built")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")


def build_hip(jobs: int = 4) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    cc = hipcc()
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        o = OBJ / (Path(src).stem + ".o")
        objs.append(o)
        cmds.append([cc, *HIPCC_FLAGS, *inc, "-c", str(CSRC / src), "-o", str(o)])
    for src in CXX_SOURCES:
        o = OBJ / (Path(src).stem + ".o")
        objs.append(o)
        cmds.append([cc, "-O2", "-fPIC", "-std=c++17", "-Wall", *inc,
                     "-D__HIP_PLATFORM_AMD__", "-c", str(CSRC / src), "-o", str(o)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, cmds))
    out = LIB / "libsplit_hip.so"
    tmp = LIB / "libsplit_hip.so.tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs),
          "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_gen() -> Path:
    LIB.mkdir(parents=True, exist_ok=True)
    out = LIB / "libsplit_gen.so"
    tmp = LIB / "libsplit_gen.so.tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          "-std=c11", "-Wall", "-Wextra", f"-I{CSRC}", f"-I{INCLUDE}",
          *[str(CSRC / s) for s in GEN_SOURCES], "-o", str(tmp), "-lm"])
    os.replace(tmp, out)
    return out


def build_dropin() -> Path:
    """The C++ caller of the drop-in headers, linked to libsplit_hip.so."""
    out = LIB / "split_dropin_sequence"
    tmp = LIB / "split_dropin_sequence.tmp"
    _run([hipcc(), "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}",
          str(ROOT / "tests" / "dropin" / "main_sequence.cpp"), f"-L{LIB}",
          "-lsplit_hip", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all(jobs: int = 4) -> list[Path]:
    return [build_gen(), build_hip(jobs), build_dropin()]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    args = ap.parse_args()
    for p in build_all(args.jobs):
        print(p)


if __name__ == "__main__":
    main()
