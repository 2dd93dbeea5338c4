"""The equal-RMSE conversion behind every learned-versus-plain figure of
bench.py: plain frames needed to reach an RMSE, from the measured plain
curve (log-log interpolation inside it, the 1/sqrt(frames) law outside it).
CPU only."""

from __future__ import annotations

import math

import numpy as np
import pytest

import bench


def _curve(fn, frames=(1, 2, 4, 8, 16, 32, 64)):
    return {k: fn(k) for k in frames}


def test_exact_on_a_pure_noise_curve():
    # RMSE = c / sqrt(N) everywhere: the answer is (c / r)^2 inside and outside the curve
    c = 0.34
    curve = _curve(lambda n: c / math.sqrt(n))
    for r in (0.5, c, 0.2, 0.1, 0.07, c / 8.0, 0.03, 0.01):
        assert bench._plain_frames_for(curve, r) == pytest.approx((c / r) ** 2, rel=1e-12)


def test_measured_points_map_to_their_frame_counts():
    rng = np.random.default_rng(2)
    vals = np.sort(rng.uniform(0.01, 1.0, size=7))[::-1]
    curve = dict(zip((1, 2, 4, 8, 16, 32, 64), vals))
    for k, r in curve.items():
        assert bench._plain_frames_for(curve, float(r)) == pytest.approx(k, rel=1e-12)


def test_log_log_interpolation_between_points():
    # between two measured points the map is a straight line in log-log space
    curve = {1: 1.0, 4: 0.6, 16: 0.25}
    r = math.sqrt(1.0 * 0.6)  # the log-midpoint of the first segment
    assert bench._plain_frames_for(curve, r) == pytest.approx(2.0, rel=1e-12)
    r2 = 0.6 ** 0.25 * 0.25 ** 0.75  # three quarters along the second segment
    assert bench._plain_frames_for(curve, r2) == pytest.approx(4.0 * 4.0 ** 0.75, rel=1e-12)


def test_monotone_and_bias_floor():
    # a curve with a bias floor (RMSE^2 = a/N + b) still maps lower RMSE to more frames
    curve = _curve(lambda n: math.sqrt(0.1 / n + 1e-4))
    rs = np.linspace(0.4, 0.005, 200)
    frames = [bench._plain_frames_for(curve, float(r)) for r in rs]
    assert all(b > a for a, b in zip(frames, frames[1:]))
    # below the last point the 1/sqrt law extends from the last measured point
    last = curve[64]
    assert bench._plain_frames_for(curve, last / 2.0) == pytest.approx(256.0, rel=1e-12)
