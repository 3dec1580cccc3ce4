"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle.cpp is rebuilt with -fsanitize=address,undefined (oracle/Makefile,
target sanitize) and the oracle's own CPU tests run again in a child Python
that has the ASan runtime preloaded and loads the sanitized library: any
out-of-bounds access, use after free or undefined behaviour (overflow,
misaligned or null access, invalid shift) aborts the run.  Leak checking is
off (the interpreter itself keeps allocations until exit).
"""

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return out if out and os.path.isabs(out) and os.path.exists(out) else None


@pytest.mark.skipif(_runtime("libasan.so") is None, reason="no ASan runtime for gcc on this host")
def test_oracle_tests_pass_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], check=True)
    env = dict(os.environ)
    env.update(LD_PRELOAD=_runtime("libasan.so"),
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               AHIP_ORACLE_LIB=str(ROOT / "oracle" / "build" / "liboracle_san.so"),
               OMP_NUM_THREADS="2")
    res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                          "tests/test_oracle.py", "tests/test_oracle_plastic.py", "tests/test_utils_cpu.py"],
                         cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    log = res.stdout[-3000:] + res.stderr[-3000:]
    assert "AddressSanitizer" not in log and "runtime error" not in log, log
    assert res.returncode == 0, log
    assert " passed" in res.stdout
