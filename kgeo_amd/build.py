"""Build the in-tree native libraries.

  libkgeo_hip.so        HIP kernels + C-ABI for gfx950 (the product)
  libkgeo_hostcheck.so  CPU instantiation of the same device headers, used
                        only by the CPU test suite and the FLOP counter
  libkgeo_hostcheck_san.so  the same CPU instantiation under AddressSanitizer
                        and UndefinedBehaviorSanitizer (gcc), for the CPU
                        suite's memory-safety run of the pass code

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


def sanitizer_runtime():
    """The AddressSanitizer runtime of the compiler build_hostcheck uses
    (clang's libclang_rt.asan, else gcc's libasan), or None."""
    cxx = host_clang()
    if cxx:
        p = subprocess.run([cxx, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                           text=True).stdout.strip()
        if p and os.path.exists(p):
            return cxx, p
    p = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    if p and os.path.exists(p):
        return "g++", p
    return None


def build_hostcheck_sanitized():
    """The CPU instantiation built with AddressSanitizer and
    UndefinedBehaviorSanitizer (host code only; nothing here is GPU code),
    with the same compiler and floating-point contract as build_hostcheck so
    that its outputs match it.  Skipped where no sanitizer runtime exists."""
    rt = sanitizer_runtime()
    if rt is None:
        print("sanitizer runtime not found: libkgeo_hostcheck_san.so not built", flush=True)
        return None
    cxx = rt[0]
    out = os.path.join(HERE, "libkgeo_hostcheck_san.so")
    tmp = out + ".tmp"
    flags = ["-O2", "-g", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unknown-pragmas", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
    if cxx != "g++":
        flags += ["-shared-libsan"]
        flags += ["-mfma", "-mavx2", "-ffp-contract=fast-honor-pragmas"] if _cpu_has_fma() else ["-ffp-contract=off"]
    else:
        flags += ["-ffp-contract=off"]
    _run([cxx] + flags + ["-o", tmp, os.path.join(CSRC, "kg_hostcheck.cpp")])
    os.replace(tmp, out)
    return out


def build_all():
    build_hostcheck()
    build_hostcheck_sanitized()
    build_opcount()
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
