"""The production random streams (rng_mode "pcg"): PCG with the RXS-M-XS
output on a 32-bit LCG state (O'Neill 2014), started at a hash of (texel,
instance, sample).  No reference counterpart; pinned here by

* the LCG recurrence's closed form s_n = a^n s_0 + c (a^n - 1) / (a - 1)
  (mod 2^32), computed without iterating, and its full period (Hull-Dobell);
* the start state, a murmur mix pinned against MurmurHash3 elsewhere;
* the output permutation being a bijection and the draws' statistics
  (uniform top byte, balanced bits, no lag-1 correlation, distinct streams
  for neighbouring texels);
* the tracer's answers that do not depend on the random numbers (the sky
  furnace) and its expectation (equal to the reference stream's mean).
"""

from __future__ import annotations

import math

import numpy as np

import oracle
from nan_amd.material import Material
from nan_amd.mesh import Mesh
from nan_amd.mesh_colors import MeshColorsLayout
from nan_amd.scene_arrays import flatten_scene
from nan_amd.scene_node import Lighting, SceneNode
from nan_amd.scenes import batch_scene
from nan_amd.transform import Transform

A, C, M32 = 747796405, 2891336453, 1 << 32


def _lcg_jump(s0: int, n: int) -> int:
    an = pow(A, n, M32 * (A - 1))
    geometric = ((an - 1) % (M32 * (A - 1))) // (A - 1)  # (a^n - 1) / (a - 1), exact before the mod
    return (pow(A, n, M32) * s0 + C * geometric) % M32


def _mix(h: int, k: int) -> int:
    return int(oracle.liboracle().orc_murmur_key_mix(h, k))


def test_state_follows_the_lcg_closed_form():
    for x, y, s in ((0, 0, 0), (12345, 7, 3), (4_000_000, 1023, 99991)):
        s0 = _mix(_mix(x, y), s)
        _, state0 = oracle.pcg_stream(x, y, s, 0)
        assert state0 == s0
        for n in (1, 2, 37, 1000):
            _, state = oracle.pcg_stream(x, y, s, n)
            assert state == _lcg_jump(s0, n), (x, y, s, n)


def test_full_period_conditions():
    # Hull-Dobell for modulus 2^32: c odd and a = 1 (mod 4)
    assert C % 2 == 1 and A % 4 == 1
    assert _lcg_jump(123, 1 << 32) == 123 and _lcg_jump(123, 1 << 31) != 123


def test_output_is_a_bijection_of_the_state():
    # 2^20 consecutive states of one stream are distinct, so are their outputs
    draws, _ = oracle.pcg_stream(5, 6, 7, 1 << 20)
    assert len(np.unique(draws)) == len(draws)


def test_draw_statistics():
    draws = np.concatenate([oracle.pcg_stream(t, t % 17, 5, 256)[0] for t in range(4096)])  # 1M draws
    n = len(draws)
    top = draws >> 24
    counts = np.bincount(top, minlength=256)
    chi2 = float(((counts - n / 256) ** 2 / (n / 256)).sum())
    assert chi2 < 255 + 5 * math.sqrt(2 * 255), chi2  # 5 sigma of chi^2(255)
    bits = ((draws[:, None] >> np.arange(32, dtype=np.uint32)) & 1).mean(axis=0)
    assert np.all(np.abs(bits - 0.5) < 5 * 0.5 / math.sqrt(n)), bits
    u = (draws >> 8).astype(np.float64) * 2.0 ** -24
    lag1 = np.corrcoef(u[:-1], u[1:])[0, 1]
    assert abs(lag1) < 5 / math.sqrt(n)
    # neighbouring texels and consecutive samples start far apart
    a = oracle.pcg_stream(1000, 3, 8, 4096)[0].astype(np.float64)
    for other in ((1001, 3, 8), (1000, 4, 8), (1000, 3, 9)):
        b = oracle.pcg_stream(*other, 4096)[0].astype(np.float64)
        assert abs(np.corrcoef(a, b)[0, 1]) < 5 / math.sqrt(4096)
        assert len(np.intersect1d(a, b)) < 4


def _ground(lighting):
    node = SceneNode()
    node.lighting = lighting
    mat = node.add_material(Material(base_color=(1.0, 1.0, 1.0), double_sided=True))
    node.add_instance(node.add_mesh(Mesh.create_quad((4.0, 4.0))), mat, node.add_transform(Transform()))
    return node


def _trace(node, rng, iterations=1, bounces=3, tpu=2.0):
    layout = MeshColorsLayout.build(node, texels_per_unit=tpu, min_resolution=1, max_resolution=8)
    sc = oracle.OracleScene(flatten_scene(node))
    sc.set_rng(rng)
    p = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    for it in range(iterations):
        sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, it, 1, bounces, it == 0,
                        threads=oracle.cpu_count())
    return layout, p


def test_sky_furnace_is_independent_of_the_stream():
    S = 0.7
    _, p = _trace(_ground(Lighting(sun_intensity=0.0, sky_radiance=(S, S, S))), "pcg")
    np.testing.assert_allclose(p[:, :3], math.pi * S, rtol=2e-6)


def test_pcg_and_reference_streams_share_the_expectation():
    # per-iteration scene means of single-iteration estimates (reset each
    # time) on a sunlit scene with occluders; the two streams' means agree to
    # their standard errors (about 0.03% of the mean: a 0.2% bias would show)
    node = batch_scene(21, 3000)
    layout = MeshColorsLayout.build(node, texels_per_unit=2.0, min_resolution=1, max_resolution=4,
                                    max_total_texels=2_000_000)
    n = layout.total_surface_texels
    per_iteration = {}
    for rng in ("reference", "pcg"):
        sc = oracle.OracleScene(flatten_scene(node))
        sc.set_rng(rng)
        p = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        m = []
        for it in range(48):
            sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, it, 1, 3, True,
                            threads=oracle.cpu_count())
            m.append(p[:n, :3].astype(np.float64).mean())
        per_iteration[rng] = np.asarray(m)
    a, b = per_iteration["reference"], per_iteration["pcg"]
    se = math.sqrt(a.var(ddof=1) / len(a) + b.var(ddof=1) / len(b))
    assert se < 2e-3 * a.mean()
    assert abs(a.mean() - b.mean()) < 5 * se, (a.mean(), b.mean(), se)
    # and the two modes are different streams
    _, x = _trace(node, "reference")
    _, y = _trace(node, "pcg")
    assert np.mean(x[:, 0] != y[:, 0]) > 0.5  # slots that see only sky agree whatever the stream
