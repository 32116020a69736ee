"""The library's host C++ (model text format, compiler and program
validation, the island search with its 16-thread breeding and its C ABI)
and the FP64 oracle, built with AddressSanitizer + UndefinedBehaviorSanitizer
and with ThreadSanitizer (tests/native/Makefile) and driven on the CPU: every
check passes and no sanitizer reports anything.  The search's evaluator is a
deterministic CPU fitness there, so two identical searches must agree
exactly, generation by generation, under ThreadSanitizer too."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
KAT = os.path.join(HERE, "golden", "kat_model.txt")
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(CLANG):
        pytest.skip("clang++ of the ROCm toolchain is not installed")
    subprocess.run(["make", "-C", NATIVE, "-j4", "-s"], check=True, timeout=600)
    return os.path.join(NATIVE, "_build")


@pytest.mark.parametrize("name,args", [("host_check_asan", [KAT]), ("host_check_tsan", [KAT]),
                                       ("oracle_check_asan", []), ("oracle_check_tsan", [])])
def test_sanitized_build_runs_clean(built, name, args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(built, name)] + args, capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    for marker in ("AddressSanitizer", "LeakSanitizer", "runtime error:", "ThreadSanitizer"):
        assert marker not in out, out[-4000:]
    assert "all checks passed" in out
