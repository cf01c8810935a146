"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` links oracle/san_driver.c with eik_oracle.c built with
-fsanitize=address,undefined; the driver calls every oracle entry point (2D/3D Cartesian with and
without refinement, 2D and 3D spherical, all whichgrad values, smoothing, the recorder, the
brute-force FMM and the bare stencil).  A sanitizer report aborts it with a non-zero status.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_runs_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(ROOT, "oracle", "san_driver")], capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    assert "all oracle entry points ran clean" in r.stdout
