"""The CPU oracle under AddressSanitizer (with leak detection) and
UndefinedBehaviorSanitizer: oracle/san_driver.c calls every oracle entry
point on rings, a plane, a lattice and the tiny sizes, built by
`make -f oracle/Makefile.san` with gcc (host code only)."""
import os
import pathlib
import shutil
import subprocess

import pytest

ORACLE = pathlib.Path(__file__).resolve().parent.parent / "oracle"


def test_oracle_under_sanitizers():
    if shutil.which("gcc") is None or shutil.which("make") is None:
        pytest.skip("gcc/make not available")
    r = subprocess.run(["make", "-s", "-f", str(ORACLE / "Makefile.san")], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(ORACLE / "_san" / "orc_san")], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "oracle sanitizer run ok" in r.stdout
