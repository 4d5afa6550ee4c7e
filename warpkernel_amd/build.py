"""Build recipe for the in-tree shared libraries (no JIT caches, no pip install).

  warpkernel_amd/libwarpkernel.so   the product: host scan, C ABI, gfx950 kernels
  warpkernel_amd/libwk_fegen.so     FE workload generators (host C++)

``python -m warpkernel_amd.build`` builds both.  hipcc compiles the HIP code for
gfx950 only; every HIP compile disables packed FP32 instructions and FMA
contraction (fusion is explicit in the sources).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Werror=shadow", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread"]

LIB_SOURCES = ["kernels.hip", "batch.hip", "dscan.hip", "assembly.hip", "colinds.hip", "noalloc.hip", "capi.cpp",
               "scan.cpp", "tune.cpp", "io.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def _compile(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    # .cpp files hold host code only; hipcc compiles them as HIP so that they
    # see the runtime headers
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", "-o", obj, path])
    return obj


def build_lib() -> str:
    with ThreadPoolExecutor(max_workers=len(LIB_SOURCES)) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    out = os.path.join(PKG, "libwarpkernel.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", out, *objs, "-pthread"])
    return out


def build_fegen() -> str:
    out = os.path.join(PKG, "libwk_fegen.so")
    _run(["g++", *HOST_FLAGS, "-shared", "-o", out, os.path.join(CSRC, "fegen.cpp")])
    return out


def build_dropin_test() -> str:
    """tests/cpp/dropin_wpk: the reference drivers' flow through include/warpkernel.hpp,
    checked against the oracle (links the product, the generators and the oracle)."""
    out_dir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "dropin_wpk")
    src = os.path.join(ROOT, "tests", "cpp", "dropin_wpk.cpp")
    _run([HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-o", out, src,
          f"-L{PKG}", "-lwarpkernel", "-lwk_fegen", f"-L{os.path.join(ROOT, 'oracle')}", "-lwk_oracle",
          "-Wl,-rpath,$ORIGIN/../../../warpkernel_amd", "-Wl,-rpath,$ORIGIN/../../../oracle"])
    return out


def build_all() -> None:
    build_lib()
    build_fegen()
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(PKG, "libwarpkernel.so"), os.path.join(PKG, "libwk_fegen.so"))
