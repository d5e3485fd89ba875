"""Build the in-tree native libraries.

  libkgeo_hip.so        HIP kernels + C-ABI for gfx950 (the product)
  libkgeo_hostcheck.so  CPU instantiation of the same device headers, used
                        only by the CPU test suite and the FLOP counter

Run:  python -m kgeo_amd.build
"""

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["kgeo_raytrace.hip", "kgeo_equatorial.hip", "kgeo_batch.hip", "kgeo_emission.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FMA fusion everywhere except code marked `#pragma clang fp contract(off)`
    # (the per-pixel setup, which must round as numpy does)
    "-ffp-contract=fast-honor-pragmas",
    "-Wno-unused-command-line-argument",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip(extra=()):
    out = os.path.join(HERE, "libkgeo_hip.so")
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = out + ".tmp"
    _run([hipcc()] + HIP_FLAGS + list(extra) + ["-I", os.path.join(ROOT, "include"), "-o", tmp] + srcs)
    os.replace(tmp, out)
    return out


def host_clang():
    for c in (shutil.which("amdclang++"), "/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    return None


def build_hostcheck():
    """CPU build of the device headers with the GPU's floating-point contract:
    clang, FMA contraction on (as hipcc does for device code), so the CPU
    suite sees the same rounding decisions (e.g. which setup code must stay
    uncontracted) as the kernels."""
    out = os.path.join(HERE, "libkgeo_hostcheck.so")
    tmp = out + ".tmp"
    cxx = host_clang()
    flags = ["-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas"]
    if cxx and _cpu_has_fma():
        flags += ["-mfma", "-mavx2", "-ffp-contract=fast-honor-pragmas"]
    else:
        cxx = "g++"
        flags += ["-ffp-contract=off"]
    _run([cxx] + flags + ["-o", tmp, os.path.join(CSRC, "kg_hostcheck.cpp")])
    os.replace(tmp, out)
    return out


def _cpu_has_fma():
    try:
        return " fma " in open("/proc/cpuinfo").read().replace("\n", " ")
    except OSError:
        return False


def build_opcount():
    out = os.path.join(HERE, "libkgeo_opcount.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas", "-ffp-contract=off",
          "-o", tmp, os.path.join(CSRC, "kg_opcount.cpp")])
    os.replace(tmp, out)
    return out


def build_all():
    build_hostcheck()
    build_opcount()
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
