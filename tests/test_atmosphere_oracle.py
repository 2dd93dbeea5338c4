"""The atmosphere oracle (oracle/atmosphere_oracle.c and the sky lookup of
oracle/nan_oracle.c) against closed forms and the invariants the tables must
hold (atmosphere.slang; scene.slang:469-518, 997-1009).  CPU only."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import atmosphere as A

G, TOP = 6.360, 6.460
RAYLEIGH = np.array([5.802, 13.558, 33.1])
OZONE = np.array([0.650, 1.881, 0.085])
MIE_SCATTER, MIE_ABSORB = 3.996, 4.4


@pytest.fixture(scope="module")
def tables():
    t = A.transmittance_lut()
    m = A.multiscattering_lut(t)
    return t, m


@pytest.mark.parametrize("alt_km", [0.0, 8.0, 25.0, 37.5, 60.0])
def test_scattering_coefficients_closed_form(alt_km):
    y = np.float32(G + alt_km / 1000.0)
    rs, mie, ext = A.scattering((0.0, y, 0.0))
    a = (float(np.float32(y) - np.float32(G))) * 1000.0  # the altitude as FP32 sees it
    want_rs = RAYLEIGH * math.exp(-a / 8.0)
    want_mie = MIE_SCATTER * math.exp(-a / 1.2)
    want_ext = want_rs + want_mie + MIE_ABSORB * math.exp(-a / 1.2) + OZONE * max(0.0, 1.0 - abs(a - 25.0) / 15.0)
    np.testing.assert_allclose(rs, want_rs, rtol=2e-6)
    assert mie == pytest.approx(want_mie, rel=2e-6, abs=1e-30)
    np.testing.assert_allclose(ext, want_ext, rtol=2e-6)


def test_sphere_distances_closed_form():
    up, side = (0.0, 1.0, 0.0), (1.0, 0.0, 0.0)
    # straight up from 1 km: the top is 99 km away
    assert A.sphere_hit((0, G + 0.001, 0), up, TOP) == pytest.approx(0.099, abs=2e-6)
    # horizontal from 1 km: the chord to the top sphere
    h = np.float32(G + 0.001)
    assert A.sphere_hit((0, h, 0), side, TOP) == pytest.approx(math.sqrt(TOP ** 2 - float(h) ** 2), rel=1e-5)
    # pointing away from a sphere we are outside of: no hit
    assert A.sphere_hit((0, 7.0, 0), up, TOP) == -1.0
    # straight down from 1 km: the ground 1 km away
    assert A.sphere_hit((0, G + 0.001, 0), (0, -1, 0), G) == pytest.approx(0.001, abs=2e-6)
    # horizontal from 1 km misses the ground
    assert A.sphere_hit((0, G + 0.001, 0), side, G) < 0.0


def _geometric_depth(h0_km: float, scale_km: float, steps: int = 40) -> float:
    """Sum of dt * exp(-alt/H) over the reference's vertical march
    (sample i at (i + 0.3)/steps of the path) as a geometric series, in km."""
    length = (TOP - G) * 1000.0 - h0_km
    d = length / steps
    r = math.exp(-d / scale_km)
    return math.exp(-h0_km / scale_km) * (0.3 * d * math.exp(-0.3 * d / scale_km)
                                          + d * math.exp(-1.3 * d / scale_km) * (1.0 - r ** (steps - 1)) / (1.0 - r))


def test_vertical_transmittance_matches_geometric_series():
    got = A.sun_transmittance((0.0, G, 0.0), (0.0, 1.0, 0.0))
    # the ozone tent integrates to 15 km exactly at these sample points
    depth = (RAYLEIGH * _geometric_depth(0.0, 8.0) + (MIE_SCATTER + MIE_ABSORB) * _geometric_depth(0.0, 1.2)
             + OZONE * 15.0) * 1e-3
    np.testing.assert_allclose(got, np.exp(-depth), rtol=5e-6)


def test_transmittance_table_horizon_and_monotonicity(tables):
    t, _ = tables
    rgb = t[..., :3].astype(np.float64)
    assert np.isfinite(rgb).all() and rgb.min() >= 0.0 and rgb.max() <= 1.0
    assert (t[..., 3] == 1.0).all()
    u = (np.arange(256) + 0.5) / 256
    v = (np.arange(64) + 0.5) / 64
    c = 2.0 * u - 1.0
    h = G + v * (TOP - G)
    # zero exactly where the ray toward the sun meets the ground
    boundary = -np.sqrt(h * h - G * G) / h
    below = c[None, :] < boundary[:, None] - 1e-4
    above = c[None, :] > boundary[:, None] + 1e-4
    assert (rgb[below] == 0.0).all()
    assert (rgb[above] > 0.0).all()
    # more light for a higher sun and for a higher point (above the horizon)
    du = np.diff(rgb, axis=1)
    ok_u = above[:, 1:] & above[:, :-1]
    assert (du[ok_u] >= -1e-6).all()
    dv = np.diff(rgb, axis=0)
    ok_v = above[1:, :] & above[:-1, :]
    assert (dv[ok_v] >= -1e-6).all()


def test_multiscattering_table_invariants(tables):
    _, m = tables
    psi = m[..., :3]
    assert np.isfinite(psi).all() and psi.min() >= 0.0
    assert (m[..., 3] == 1.0).all()
    # sun above the horizon lights the air: every column with u > 0.6 is positive at the ground
    assert (psi[0, 20:] > 0.0).all()


@pytest.mark.parametrize("sun", [(0.0, 0.5, 0.866), (0.3, 0.2, -0.9)])
def test_sky_view_table_is_mirror_symmetric_about_the_sun(tables, sun):
    t, m = tables
    s = A.sky_view_lut(t, m, sun)
    rgb = s[..., :3]
    assert np.isfinite(rgb).all() and rgb.min() >= 0.0
    np.testing.assert_allclose(rgb, rgb[:, ::-1], rtol=2e-5, atol=1e-9)
    # toward the sun (middle column) is brighter than away from it near the horizon
    row = 100
    assert (rgb[row, 100] > rgb[row, 0]).all()


def _texel_directions(sun):
    """Directions through every texel centre of the sky-view table, built
    with the generator's altitude/azimuth mapping (atmosphere.slang:359-380)
    in float64 and rotated so the sun's azimuth is the table's azimuth 0."""
    u = (np.arange(200) + 0.5) / 200
    v = (np.arange(200) + 0.5) / 200
    az = (u - 0.5) * 2.0 * np.pi
    adj = np.where(v < 0.5, -(1.0 - 2.0 * v) ** 2, (2.0 * v - 1.0) ** 2)
    h = G + 0.0002
    horizon = math.acos(math.sqrt(h * h - G * G) / h) - 0.5 * math.pi
    alt = adj * 0.5 * math.pi - horizon
    AL, AZ = np.meshgrid(alt, az, indexing="ij")
    local = np.stack([np.cos(AL) * np.sin(AZ), np.sin(AL), -np.cos(AL) * np.cos(AZ)], axis=-1)
    # table frame: forward (azimuth 0) = -z, right = +x; scene frame: forward =
    # the sun's horizontal direction, right = cross(sun, up)
    s = np.asarray(sun, dtype=np.float64)
    fwd = np.array([s[0], 0.0, s[2]])
    fwd /= np.linalg.norm(fwd)
    right = np.cross(s, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    world = local[..., 0:1] * right + local[..., 1:2] * np.array([0.0, 1.0, 0.0]) - local[..., 2:3] * fwd
    return world.reshape(-1, 3), alt


@pytest.mark.parametrize("sun", [(0.0, 0.5, -0.866), (0.0, 0.5, 0.866), (0.6, 0.3, 0.2)])
def test_sky_lookup_inverts_the_table_mapping(tables, sun):
    from nan_amd.scene_node import Lighting

    t, m = tables
    rec = Lighting(sun_direction=sun, sun_intensity=2.0, sun_color=(1.0, 0.5, 0.25)).pack()
    sky = A.sky_view_lut(t, m, rec[0:3])
    dirs, alt = _texel_directions(rec[0:3])
    got = A.sky_view_lookup(sky, rec, dirs)[:, :3].reshape(200, 200, 3)
    want = sky[..., :3] * np.array([2.0, 1.0, 0.5], dtype=np.float32)
    # the lookup clamps the altitude to +-(pi/2 - 0.01): rows past it do not map back
    rows = np.abs(alt) < 0.5 * math.pi - 0.011
    scale = float(want.max())
    np.testing.assert_allclose(got[rows], want[rows], rtol=0, atol=2e-4 * scale)


def test_sky_lookup_edge_directions(tables):
    from nan_amd.scene_node import Lighting

    t, m = tables
    rec = Lighting.reference_default().pack()
    sky = A.sky_view_lut(t, m, rec[0:3])
    dirs = np.array([[0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 0, 0], [np.nan, 0, 0], [0, 0, 0]],
                    dtype=np.float32)
    out = A.sky_view_lookup(sky, rec, dirs)
    assert np.isfinite(out[:5]).all() and (out[:5, :3] >= 0.0).all()
    # toward the sun's azimuth at the horizon is brighter than away from it
    assert (out[2, :3] > out[3, :3]).all()


def test_view_transmittance_reads_the_bottom_row(tables):
    t, _ = tables
    sun = np.array([0.0, 0.5, 0.866])
    got = A.view_transmittance(t, sun)
    c = float(np.float32(sun[1] / np.linalg.norm(sun)))
    x = (0.5 + 0.5 * c) * 256 - 0.5
    x0 = int(math.floor(x))
    w = x - x0
    # the view point (0.2 km) lies below the first row's centre: clamp to row 0
    want = t[0, x0, :3] * (1 - w) + t[0, x0 + 1, :3] * w
    np.testing.assert_allclose(got, want, rtol=1e-5)


def test_multiscattering_table_is_linear_in_the_transmittance_table(tables):
    # every sun term and the ground bounce carry one transmittance factor and
    # f_ms carries none: doubling the table doubles the result, bit for bit
    # (a factor of two commutes with FP32 rounding and with the bilinear read)
    t, m = tables
    m2 = A.multiscattering_lut(2.0 * t)
    np.testing.assert_array_equal(m2[..., :3], 2.0 * m[..., :3])
    # a table of zeros leaves nothing: no term without a sun transmittance
    assert (A.multiscattering_lut(np.zeros_like(t))[..., :3] == 0.0).all()


def test_multiscattering_is_zero_with_the_sun_under_the_ground(tables):
    # ground level, sun straight down: every sample point has the planet
    # between it and the sun, and the ground bounce needs dot(pos, sun) > 0
    _, m = tables
    assert (m[0, 0, :3] == 0.0).all()
    assert (m[0, -1, :3] > 0.0).all()


@pytest.mark.parametrize("sun", [(0.0, 0.5, 0.866), (0.3, 0.2, -0.9)])
def test_sky_view_table_is_linear_in_its_two_tables(tables, sun):
    t, m = tables
    s = A.sky_view_lut(t, m, sun)
    # both tables doubled: the radiance doubles bit for bit
    np.testing.assert_array_equal(A.sky_view_lut(2.0 * t, 2.0 * m, sun)[..., :3], 2.0 * s[..., :3])
    # single scattering (sun transmittance) plus multiple scattering (psi); the
    # in-segment integral (in - in * T) / ext cancels in FP32 on optically thin
    # steps (relative error ~1e-7 / (1 - T)), so the split sums agree to 1e-3
    single = A.sky_view_lut(t, np.zeros_like(m), sun)[..., :3].astype(np.float64)
    multi = A.sky_view_lut(np.zeros_like(t), m, sun)[..., :3].astype(np.float64)
    np.testing.assert_allclose(single + multi, s[..., :3], rtol=1e-3, atol=1e-12)
    assert (single > 0.0).mean() > 0.4 and (multi > 0.0).all()
