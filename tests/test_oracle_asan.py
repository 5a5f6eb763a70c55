"""The oracle's CPU tests under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle asan` builds the same source with -fsanitize=address,undefined
(and no recovery: the first report aborts); a child interpreter with the ASan
runtime preloaded runs the oracle test files against that build.  Any report
fails the run.
"""

import os
import subprocess
import sys

import pytest

from .conftest import ROOT

FILES = ["tests/test_oracle_primitives.py", "tests/test_oracle_fixtures.py", "tests/test_oracle_unitigs.py",
         "tests/test_oracle_merge_pins.py", "tests/test_oracle_clean.py", "tests/test_oracle_bubble.py",
         "tests/test_oracle_stats.py", "tests/test_formats.py"]


def test_oracle_tests_clean_under_sanitizers():
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir, "asan"], check=True)
    def gcc_lib(name):
        return subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    asan, stdcxx = gcc_lib("libasan.so"), gcc_lib("libstdc++.so")
    if not all(os.path.isabs(p) and os.path.exists(p) for p in (asan, stdcxx)):
        pytest.skip("libasan is not installed")
    env = dict(os.environ)
    # the ASan runtime must come first, and libstdc++ must be loaded when it
    # starts (its __cxa_throw interceptor binds then); anything already
    # preloaded stays
    env["LD_PRELOAD"] = asan + ":" + stdcxx + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else "")
    env["ORACLE_LIB"] = os.path.join(odir, "liboracle_asan.so")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"  # the interpreter itself keeps objects alive at exit
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        *FILES], cwd=ROOT, env=env, capture_output=True, text=True, timeout=1200)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
