"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(make -C oracle sanitize) runs its grid set-up, line solves, Poisson solve,
tg substeps, moments and spline lookups on small grids without a report."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_under_asan_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("no gcc")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    r = subprocess.run([os.path.join(ROOT, "oracle", "_san", "sanitize_check")], capture_output=True, text=True,
                       timeout=300, env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert r.returncode == 0, r.stderr[-2000:]
    assert "sanitize_check: done" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
