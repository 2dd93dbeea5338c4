"""The oracle's tone map (oracle/nan_oracle.c, after tone_mapper.slang:8-90)
against what the mathematics of the reference's curve fixes.  CPU only.

With contrast 1 the LPM luma curve is f(L) = L / (L X + Y), X and Y solved
so that the mid grey midIn = 12 * 0.18 * 2^-exposure maps to 0.18 and the
HDR maximum 12 maps to 1.  Grey inputs pass through the crosstalk steps
unchanged (equal channels, luma weights summing to one), so their output is
Rec709(min(1, f(g))) per channel; a coloured input that the first step does
not clip keeps its hue and has output luma f(L) (the mapper preserves luma)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle

LUMA = np.array([0.2126, 0.7152, 0.0722])


def _rec709(x):
    return 4.5 * x if x < 0.018 else 1.099 * x ** 0.45 - 0.099


def _curve(L, exposure):
    mid_in = 12.0 * 0.18 * 2.0 ** -exposure
    X = (2.16 - mid_in) / (2.16 - 0.18 * mid_in)
    Y = 12.0 * (1.0 - X)
    return L / (L * X + Y)


def _tonemap(rgb, exposure):
    inp = np.zeros((len(rgb), 4), dtype=np.float32)
    inp[:, :3] = rgb
    return oracle.tonemap(inp, exposure)


@pytest.mark.parametrize("exposure", [-2.0, 0.0, 0.5, 1.5])
def test_grey_known_answers(exposure):
    mid_in = 12.0 * 0.18 * 2.0 ** -exposure
    greys = [0.0, mid_in, 12.0, 30.0, 1e-3, 0.05, 0.7, 5.0]
    out = _tonemap(np.array([[g, g, g] for g in greys]), exposure)
    assert (out[:, 3] == 1.0).all()
    # zero stays zero; the mid grey maps to Rec709(0.18); 12 maps to 1
    np.testing.assert_array_equal(out[0, :3], 0.0)
    np.testing.assert_allclose(out[1, :3], _rec709(0.18), rtol=2e-6)
    np.testing.assert_allclose(out[2, :3], 1.0, rtol=2e-6)
    X = (2.16 - mid_in) / (2.16 - 0.18 * mid_in)
    if X >= 0.0:  # the curve rises to 1 at 12 and keeps rising: above 12 saturates
        np.testing.assert_allclose(out[3, :3], 1.0, rtol=2e-6)
    else:  # mid grey above 2.16: the curve has a pole at -Y/X past 12 and is negative beyond
        assert -12.0 * (1.0 - X) / X < 30.0
        np.testing.assert_array_equal(out[3, :3], 0.0)
    for g, o in zip(greys[4:], out[4:]):
        want = _rec709(min(1.0, _curve(g, exposure)))
        np.testing.assert_allclose(o[:3], want, rtol=1e-5, atol=1e-7)


def test_curve_is_monotone_and_bounded():
    g = np.geomspace(1e-5, 100.0, 400)
    out = _tonemap(np.stack([g, g, g], axis=1), 0.0)[:, 0]
    assert (np.diff(out) >= 0).all() and out.min() >= 0.0 and out.max() <= 1.0


@pytest.mark.parametrize("exposure", [0.0, 1.0])
def test_unclipped_colours_keep_hue_and_luma(exposure):
    rng = np.random.default_rng(5)
    rgb = rng.uniform(0.0, 1.0, size=(2000, 3)) * rng.uniform(0.01, 3.0, size=(2000, 1))
    out = _tonemap(rgb, exposure)[:, :3].astype(np.float64)
    L = rgb @ LUMA
    f = _curve(L, exposure)
    m = rgb.max(axis=1)
    s = f / ((rgb / m[:, None]) @ LUMA)  # the first step's scale of the hue ratio
    ok = (s < 0.999) & (m > 0)
    assert ok.sum() > 500
    # invert Rec709 per channel to reach the linear mapper output
    lin = np.where(out < 4.5 * 0.018, out / 4.5, ((out + 0.099) / 1.099) ** (1.0 / 0.45))
    np.testing.assert_allclose(lin[ok] @ LUMA, f[ok], rtol=2e-5, atol=1e-7)
    np.testing.assert_allclose(lin[ok], (rgb / m[:, None])[ok] * s[ok, None], rtol=5e-5, atol=1e-6)


def test_saturated_colour_moves_toward_white_with_its_luma():
    # pure red brighter than the first step can carry (at exposure 0 the curve
    # is L / 12, so red clips above 12): the crosstalk term adds green and blue
    # until the output luma reaches f(L)
    exposure = 0.0
    r = 30.0
    out = _tonemap(np.array([[r, 0.0, 0.0]]), exposure)[0, :3].astype(np.float64)
    lin = np.array([(o / 4.5) if o < 4.5 * 0.018 else ((o + 0.099) / 1.099) ** (1.0 / 0.45) for o in out])
    f = _curve(0.2126 * r, exposure)
    assert lin[0] == pytest.approx(1.0, abs=1e-6)
    assert lin[1] > 0.0 and lin[2] > 0.0
    # crosstalk (1, 0.5, 0.0625): green gets 8x the blue share of the added light
    assert lin[1] / lin[2] == pytest.approx(8.0, rel=1e-4)
    assert float(lin @ LUMA) == pytest.approx(f, rel=2e-5)


def test_negative_and_black_inputs_give_black():
    out = _tonemap(np.array([[-1.0, -2.0, -0.5], [0.0, 0.0, 0.0]]), 0.0)
    np.testing.assert_array_equal(out[:, :3], 0.0)
    assert math.isclose(float(out[0, 3]), 1.0)
