"""Build the native libraries in-tree.

  libocu_hip.so  hand-written HIP kernels for gfx950 + their C ABI (include/ocu_hip.h)
  libocu.so      host C++ drop-in classes (namespace ocu) + handle C ABI (include/ocu.h),
                 compiled by g++ with no HIP headers, linked against libocu_hip.so

Run ``python -m opencurrent_amd.build`` (or __graft_entry__.build()).
"""
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
LIBDIR = PKG / "lib"
OBJDIR = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OCU_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no FMA contraction, so every kernel performs exactly the
# operations it spells out (bit-comparable with the CPU restatement).
# -packed-fp32-ops: no packed FP32 instructions are generated.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(map(str, cmd))}\n{p.stdout}\n{p.stderr}")
    return p


def build_hip(jobs=8, verbose=False):
    OBJDIR.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted((PKG / "csrc" / "hip").glob("*.hip"))
    objs = []

    def comp(src):
        obj = OBJDIR / (src.stem + ".hip.o")
        _run([HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(comp, srcs))
    out = LIBDIR / "libocu_hip.so"
    tmp = LIBDIR / "libocu_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_host(jobs=8, verbose=False):
    OBJDIR.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    srcs = sorted((PKG / "csrc" / "host").glob("*.cpp"))
    if not srcs:
        return None

    def comp(src):
        obj = OBJDIR / (src.stem + ".o")
        _run(["g++", *HOST_FLAGS, "-I", str(ROOT / "include"), "-I", str(PKG / "csrc" / "host" / "include"),
              "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(comp, srcs))
    out = LIBDIR / "libocu.so"
    tmp = LIBDIR / "libocu.so.tmp"
    _run(["g++", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
          "-L", str(LIBDIR), "-locu_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_all(verbose=False):
    build_hip(verbose=verbose)
    build_host(verbose=verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
