"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a
standalone driver (oracle/asan_check.c) built with both sanitizers runs every
exported oracle routine on a small seeded system, including a reference pair
with a zero-overlap orbital, and checks the corrected Wick elements against
the generalised Slater-Condon route.  CPU only."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None and shutil.which("cc") is None, reason="no C compiler")
def test_oracle_under_sanitizers():
    b = subprocess.run(["make", "-C", ORACLE, "asan"], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(ORACLE, "asan_check")], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout, r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-3000:]
    assert "asan_check ok" in r.stdout
