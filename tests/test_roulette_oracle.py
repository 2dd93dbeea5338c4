"""Russian roulette in the oracle's path loop (an extension: the reference's
paths always run to max_bounces, texture_space_path_tracer.slang:57-84).

Pinned by a closed-form scene: the inside of a closed box whose six
single-sided walls all emit Le, with albedo a and no sun or sky.  Without
roulette every path sees Le on each of its n hits, weighted by a^j, so every
texel's irradiance is exactly pi * Le * (1 - a^n) / (1 - a).  With roulette
and a grey albedo the survival probability is a and the reweighted throughput
is exactly 1, so one sample is pi * Le * m for an integer m in 1..n with
P(m > j) = a^j: a dropped reweighting, a wrong survival probability or a draw
on the last bounce fails one of the checks below.  CPU only."""

from __future__ import annotations

import math

import numpy as np

import oracle
from nan_amd.material import Material
from nan_amd.mesh import Mesh
from nan_amd.mesh_colors import MeshColorsLayout
from nan_amd.scene_arrays import flatten_scene
from nan_amd.scene_node import Lighting, SceneNode
from nan_amd.transform import Transform


def _emitting_box(albedo, Le: float = 1.0) -> SceneNode:
    """Closed box [-1,1] x [0,2] x [-1,1]; every wall faces inward and emits Le."""
    node = SceneNode()
    node.lighting = Lighting.black()
    mat = node.add_material(Material(base_color=tuple(albedo), emissive=(Le, Le, Le)))
    floor = node.add_mesh(Mesh.create_quad((2, 2)))
    node.add_instance(floor, mat, node.add_transform(Transform((0, 0, 0))))
    node.add_instance(floor, mat, node.add_transform(Transform((0, 2, 0), (math.pi, 0, 0))))
    node.add_instance(node.add_mesh(Mesh.create_quad_xy((2, 2), face_positive_z=True)), mat,
                      node.add_transform(Transform((0, 1, -1))))
    node.add_instance(node.add_mesh(Mesh.create_quad_xy((2, 2), face_positive_z=False)), mat,
                      node.add_transform(Transform((0, 1, 1))))
    node.add_instance(node.add_mesh(Mesh.create_quad_yz((2, 2), face_positive_x=True)), mat,
                      node.add_transform(Transform((-1, 1, 0))))
    node.add_instance(node.add_mesh(Mesh.create_quad_yz((2, 2), face_positive_x=False)), mat,
                      node.add_transform(Transform((1, 1, 0))))
    return node


def _frames(node, roulette: bool, bounces: int, frames: int, tpu: float = 24.0):
    """Per-frame irradiance of every surface texel, float32[frames, N, 3]."""
    node.russian_roulette = roulette
    layout = MeshColorsLayout.build(node, texels_per_unit=tpu, min_resolution=1, max_resolution=64)
    sc = oracle.OracleScene(flatten_scene(node))
    out = []
    for it in range(frames):
        payload = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, payload, it, 1, bounces,
                        True, threads=oracle.cpu_count())
        out.append(payload[: layout.total_surface_texels, :3].copy())
    return np.stack(out)


def test_without_roulette_every_texel_sees_the_geometric_sum():
    a, n = 0.5, 4
    e = _frames(_emitting_box((a, a, a)), False, n, 1)
    want = math.pi * (1.0 - a ** n) / (1.0 - a)
    np.testing.assert_allclose(e, want, rtol=2e-5)


def test_roulette_gives_integer_hit_counts_with_geometric_tail():
    a, n = 0.5, 4
    e = _frames(_emitting_box((a, a, a)), True, n, 2)
    m = e[..., 0] / math.pi
    k = np.rint(m)
    # reweighted throughput is exactly 1: every sample is pi * Le * (hits)
    np.testing.assert_allclose(m, k, atol=1e-5)
    np.testing.assert_array_equal(e[..., 0], e[..., 1])
    np.testing.assert_array_equal(e[..., 0], e[..., 2])
    assert k.min() == 1 and k.max() == n
    # survival probability a per bounce: P(m > j) = a^j for j < n
    count = k.size
    for j in range(1, n):
        frac = float((k > j).mean())
        sigma = math.sqrt(a ** j * (1.0 - a ** j) / count)
        assert abs(frac - a ** j) < 5.0 * sigma, (j, frac, a ** j)
    # the mean is the no-roulette value
    want = (1.0 - a ** n) / (1.0 - a)
    assert abs(k.mean() - want) < 5.0 * k.std() / math.sqrt(count)


def test_roulette_is_unbiased_for_coloured_albedo():
    albedo = (0.8, 0.5, 0.2)
    n = 5
    e = _frames(_emitting_box(albedo), True, n, 6)
    flat = e.reshape(-1, 3)
    for c, a in enumerate(albedo):
        want = math.pi * (1.0 - a ** n) / (1.0 - a)
        err = 5.0 * flat[:, c].std() / math.sqrt(flat.shape[0])
        assert abs(flat[:, c].mean() - want) < err, (c, flat[:, c].mean(), want, err)
    # paths do end early: some samples are below the full-length value in every channel
    full = math.pi * (1.0 - albedo[0] ** n) / (1.0 - albedo[0])
    assert (flat[:, 0] < 0.9 * full).mean() > 0.2


def test_one_bounce_paths_are_unchanged_by_roulette():
    # no path can take another bounce, so roulette draws nothing
    node = _emitting_box((0.6, 0.3, 0.9))
    node.lighting = Lighting.sun_and_sky(40.0, 20.0, sky=0.2)
    plain = _frames(node, False, 1, 2, tpu=8.0)
    rr = _frames(node, True, 1, 2, tpu=8.0)
    np.testing.assert_array_equal(plain.view(np.uint32), rr.view(np.uint32))
