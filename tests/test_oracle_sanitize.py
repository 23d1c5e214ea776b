"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

oracle/sanitize_check.c drives every entry point of oracle/dcrf_oracle.h that the tests
use (2-D and generic models, the three compatibility kinds, the four normalisations, the
three kernel types, d = 1..7, L = 1, 2, 3 and 21, odd sizes, the three objectives, the
kernel gradient, parameter get/set, the FP64 re-run).  Built by `make -C oracle
sanitize` with gcc -fsanitize=address,undefined; a sanitizer report (an out-of-bounds
access, a leak, undefined behaviour) makes the run fail.
"""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_under_sanitizers(tmp_path):
    out = tmp_path / "sanitize_check"
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize", "SAN_OUT=%s" % out], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(out)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.stdout + r.stderr)[-6000:]
