"""The host code of the library and the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (no GPU).

tests/sanitize/Makefile builds the library's sources and oracle/orc.c with both
sanitizers (host code only; device code as for the product) into one program,
tests/sanitize/host_checks.cpp, which exercises the layout, the planner against
the oracle's, dispatcher contract violations, argument checks, batch work-list
planning (dependency order and item coverage), TGA round trips with truncated
and mutated files, and the oracle's emulators on odd sizes.  Leak detection is
on.  Any sanitizer report fails the run.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.mark.skipif(not (os.path.exists(os.path.join(ROCM, "bin", "hipcc"))
                         and os.path.exists(os.path.join(ROCM, "llvm", "bin", "clang++"))
                         and shutil.which("make")),
                    reason="needs hipcc, ROCm's clang and make")
def test_host_code_is_clean_under_asan_and_ubsan():
    jobs = str(max(1, min(8, os.cpu_count() or 1)))
    build = subprocess.run(["make", "-C", HERE, "-j", jobs, f"ROCM={ROCM}"], capture_output=True, text=True)
    assert build.returncode == 0, build.stdout[-4000:] + build.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    run = subprocess.run([os.path.join(HERE, "_build", "host_checks")], capture_output=True, text=True, env=env,
                         timeout=600)
    assert run.returncode == 0, run.stdout[-4000:] + run.stderr[-8000:]
    assert "all host checks passed" in run.stdout
    assert "runtime error" not in run.stderr  # UBSan reports, should any be non-fatal
