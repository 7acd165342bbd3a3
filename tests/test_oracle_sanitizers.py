"""The oracle's pins rerun on a build of the oracle under AddressSanitizer and
UndefinedBehaviorSanitizer (oracle/Makefile target `sanitize`, every report
fatal), in a child process with the ASan runtime preloaded: an out-of-bounds
access, a use after free or undefined arithmetic in the checker fails here
instead of silently shifting the numbers it checks against."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "oracle", "_san", "librapt_oracle_san.so")


def _runtime(name):
    try:
        out = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True,
                             text=True, timeout=30).stdout.strip()
    except (OSError, subprocess.SubprocessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


def test_oracle_pins_pass_under_sanitizers():
    if not os.path.exists(SAN):
        pytest.skip("sanitized oracle not built (make -C oracle sanitize)")
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("no ASan runtime")
    env = dict(os.environ)
    env["RAPT_ORACLE_LIB"] = SAN
    # keep whatever the environment already preloads after the ASan runtime
    env["LD_PRELOAD"] = ":".join(p for p in (asan, os.environ.get("LD_PRELOAD", "")) if p)
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", "tests/test_oracle.py",
                        "tests/test_oracle_switching.py", "tests/test_oracle_equatorial.py",
                        "tests/test_userfields.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "ERROR: AddressSanitizer" not in tail and \
        "runtime error" not in tail, tail
