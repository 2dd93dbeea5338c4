"""Pins of the oracle's Mesh Colors lookups by what the mathematics fixes.

* The resolve's 3-tap lookup (mesh_colors_sample_irradiance,
  mesh_colors.slang:365-426) is barycentric interpolation over the lattice
  sub-triangle that holds the query.  Interpolating a linear field
  reproduces it, so a cache whose every slot holds an affine function of its
  own lattice barycentrics must return that function of the query's
  barycentrics, up to FP32 rounding.  A swapped texel/weight pair, a wrong
  upper/lower sub-triangle choice or a transposed (i, j) fails this.
* A query exactly on a lattice point returns that texel's value bit for bit
  (front and back slots, float payload and RGB9E5 words).
* mesh_colors_nearest_texel (mesh_colors.slang:325-363) picks the sub-
  triangle corner with the largest barycentric weight, which is the lattice
  point nearest the query in the L1 distance of the three barycentric
  coordinates (lattice units).  Checked against a brute-force search over all
  (R+1)(R+2)/2 lattice points, away from near-ties (the reference breaks ties
  within 1e-3 in a fixed order).
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle


def count_of(R: int) -> int:
    return (R + 1) * (R + 2) // 2


def _rows(R: int, texel_offset: int, back_offset: int):
    count = (R + 1) * (R + 2) // 2
    inst = np.array([0, 1, texel_offset, count], dtype=np.uint32)
    face = np.array([0, R], dtype=np.uint32)
    side = np.array([back_offset, 1], dtype=np.uint32)
    return inst, face, side, count


def _lattice(R: int):
    ij = [(i, j) for i in range(R + 1) for j in range(R + 1 - i)]
    addr = np.array([i * (R + 1) - i * (i - 1) // 2 + j for i, j in ij])
    assert (addr == np.arange(len(ij))).all()  # row-major order of the lattice
    return np.array(ij, dtype=np.int64)


def _random_barycentrics(rng, n):
    u = rng.random((n, 2))
    flip = u.sum(axis=1) > 1.0
    u[flip] = 1.0 - u[flip]
    b = np.zeros((n, 3), dtype=np.float32)
    b[:, 0] = u[:, 0]
    b[:, 1] = u[:, 1]
    b[:, 2] = np.float32(1.0) - b[:, 0] - b[:, 1]
    return b


@pytest.mark.parametrize("R", [1, 2, 4, 8, 16, 64])
def test_lookup_reproduces_a_linear_field(R):
    rng = np.random.default_rng(R)
    T0, gap = 7, 3
    inst, face, side, count = _rows(R, T0, T0 + count_of(R) + gap)
    ij = _lattice(R).astype(np.float64)
    payload = np.zeros((T0 + 2 * count + gap, 4), dtype=np.float32)
    # front: the lattice barycentrics themselves; back: another affine field
    front = np.stack([ij[:, 0] / R, ij[:, 1] / R, 1.0 - (ij[:, 0] + ij[:, 1]) / R], axis=1)
    back = np.stack([2.0 * ij[:, 0] / R + 1.0, 3.0 * ij[:, 1] / R - 1.0, 0.5 - ij[:, 0] / R], axis=1)
    payload[T0:T0 + count, :3] = front
    payload[side[0]:side[0] + count, :3] = back
    # FP32 rounding: the three fractional weights come from b * R, each within
    # ulp(R)/2 of exact, so the result is within about 4 R 2^-24 max|field|;
    # a swapped tap misses by about 1/(2R)
    tol = 4.0 * R * 2.0 ** -24 + 1e-6
    bs = _random_barycentrics(rng, 400)
    for b in bs:
        bd = b.astype(np.float64)
        got_f = oracle.sample_irradiance(b, inst, face, side, back=False, payload=payload)
        np.testing.assert_allclose(got_f, bd, rtol=0, atol=tol)
        got_b = oracle.sample_irradiance(b, inst, face, side, back=True, payload=payload)
        want_b = np.array([2.0 * bd[0] + 1.0, 3.0 * bd[1] - 1.0, 0.5 - bd[0]])
        np.testing.assert_allclose(got_b, want_b, rtol=0, atol=3.0 * tol)


@pytest.mark.parametrize("R", [1, 2, 4, 8, 16, 64])
def test_lookup_at_lattice_points_returns_the_texel(R):
    rng = np.random.default_rng(100 + R)
    T0 = 5
    count = count_of(R)
    inst, face, side, _ = _rows(R, T0, T0 + count)
    payload = np.zeros((T0 + 2 * count, 4), dtype=np.float32)
    payload[:, :3] = rng.uniform(0.0, 50.0, size=(len(payload), 3)).astype(np.float32)
    words = oracle.pack_rgb9e5(payload[:, :3])
    for k, (i, j) in enumerate(_lattice(R)):
        b = np.array([i / R, j / R, (R - i - j) / R], dtype=np.float32)  # exact for R a power of two
        for back, slot in ((False, T0 + k), (True, T0 + count + k)):
            got = oracle.sample_irradiance(b, inst, face, side, back=back, payload=payload)
            np.testing.assert_array_equal(got.view(np.uint32), payload[slot, :3].view(np.uint32))
            got = oracle.sample_irradiance(b, inst, face, side, back=back, packed=words)
            np.testing.assert_array_equal(got, oracle.unpack_rgb9e5(int(words[slot]))[:3])


@pytest.mark.parametrize("R", [1, 2, 4, 8, 16])
def test_nearest_texel_is_the_l1_nearest_lattice_point(R):
    rng = np.random.default_rng(200 + R)
    inst, face, _, count = _rows(R, 0, 0)
    lat = _lattice(R).astype(np.float64)
    pts = np.stack([lat[:, 0], lat[:, 1], R - lat[:, 0] - lat[:, 1]], axis=1)
    checked = upper = 0
    for b in _random_barycentrics(rng, 3000):
        sc = b.astype(np.float64) * R
        d = np.abs(pts - sc).sum(axis=1)
        order = np.argsort(d)
        if d[order[1]] - d[order[0]] < 1e-2:  # near-tie: the reference's fixed order decides
            continue
        frac = sc - np.floor(sc)
        if np.any(np.minimum(frac, 1.0 - frac) < 1e-5):  # on a sub-triangle edge
            continue
        upper += abs(frac.sum() - 2.0) < 1e-3
        assert oracle.nearest_texel(b, inst, face) == int(order[0]), (R, b, sc)
        checked += 1
    assert checked > 2000
    if R > 1:
        assert upper > 300  # both sub-triangle orientations are exercised
