"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): libnan_hip's BVH builder (nan_bvh.cpp) and the oracle, driven by
oracle/sanitize_check.cpp over seeded random scenes (degenerate and
axis-aligned triangles, mirrored instances, axis-parallel rays).  The driver
also checks the builder's invariants and that the oracle's BVH equals its
brute force."""

from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_bvh_builder_and_oracle_under_asan_ubsan():
    build = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], capture_output=True, text=True)
    if build.returncode != 0 and ("asan" in build.stderr or "ubsan" in build.stderr):
        pytest.skip("compiler without sanitizer runtime: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(ROOT / "oracle" / "_build" / "sanitize_check")], capture_output=True, text=True,
                         timeout=600)
    assert run.returncode == 0, (run.stdout + run.stderr)[-4000:]
    assert "sanitize_check ok" in run.stdout
