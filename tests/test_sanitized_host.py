"""Memory safety of the pass code: the CPU instantiation of the device
source built with AddressSanitizer and UndefinedBehaviorSanitizer runs every
pass (heads, far, repair, walk, fixup; theta and radial) over ragged pixel
counts, 3..257 samples, 1..7 chunks per pixel and five regimes, with and
without t and phi (NULL outputs), plus the count formulas; outputs must match
the plain CPU build.  The sanitizer runtime is preloaded into a child
process (tests/sanitized_run.py)."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "kgeo_amd", "libkgeo_hostcheck_san.so")


def _runtime():
    from kgeo_amd.build import sanitizer_runtime
    rt = sanitizer_runtime()
    return rt[1] if rt else None


@pytest.mark.skipif(not os.path.exists(SAN) or _runtime() is None, reason="sanitizer build or runtime missing")
def test_passes_under_address_and_ub_sanitizers():
    env = dict(os.environ, LD_PRELOAD=_runtime(), ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", PYTHONPATH=ROOT)
    p = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitized_run.py")], env=env,
                       capture_output=True, text=True, timeout=900)
    report = (p.stdout + p.stderr)[-4000:]
    assert p.returncode == 0, report
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, report
    assert "sanitized runs clean" in p.stdout, report
