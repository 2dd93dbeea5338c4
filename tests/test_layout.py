"""Mesh Colors layout and lattice math: product vs oracle vs the reference's
known answers (test_mesh_colors.py:17-163)."""

from __future__ import annotations

import numpy as np
import pytest

from nan_amd.material import MATERIAL_FLAG_DOUBLE_SIDED, Material
from nan_amd.mesh_colors import (
    MESH_COLORS_INVALID_OFFSET,
    MESH_COLORS_SIDE_FLAG_DOUBLE_SIDED,
    MeshColorsLayout,
    address_to_ij,
    barycentric_at,
    colors_per_patch,
    ij_to_address,
)
from nan_amd.scene_node import SceneNode
from nan_amd.scenes import batch_scene, cornell_box, synthetic_scene
import oracle
from oracle import layout_oracle


def test_address_round_trip():
    # test_mesh_colors.py:17-24
    for resolution in (1, 2, 4, 8, 16):
        for address in range(colors_per_patch(resolution)):
            i, j = address_to_ij(address, resolution)
            assert ij_to_address(i, j, resolution) == address
            bary = barycentric_at(i, j, resolution)
            assert sum(bary) == pytest.approx(1.0)
            assert all(v >= 0.0 for v in bary)


def test_address_forms_agree_exhaustively():
    # closed form with fix-up (mesh_colors.py:32-52) == bisection (mesh_colors.slang:48-65)
    for resolution in range(1, 65):
        for address in range(colors_per_patch(resolution)):
            a = layout_oracle.address_to_ij_sqrt(address, resolution)
            assert address_to_ij(address, resolution) == a
            assert oracle.address_to_ij(address, resolution) == a


def test_colors_per_patch_closed_form():
    for r in range(1, 129):
        count = sum(1 for i in range(r + 1) for j in range(r + 1 - i))
        assert colors_per_patch(r) == count


def test_demo_layout_offsets_and_resolution_clamp():
    # test_mesh_colors.py:26-82
    node = SceneNode.demo()
    layout = MeshColorsLayout.build(node, texels_per_unit=64.0, min_resolution=2, max_resolution=8,
                                    max_total_texels=100_000)
    assert layout.total_texel_count > 0
    assert len(layout.instance_infos) == len(node.instances)
    assert len(layout.face_infos) == sum(node.meshes[m].triangle_count for m, _, _ in node.instances)
    expected_payload = 0
    expected_surface = 0
    ranges = []
    for inst, side in zip(layout.instance_infos, layout.side_infos):
        assert inst.texel_offset == expected_payload
        local = 0
        for f in range(inst.face_count):
            face = layout.face_infos[inst.face_offset + f]
            assert face.address == local
            assert 2 <= face.resolution <= 8
            local += colors_per_patch(face.resolution)
        assert inst.texel_count == local
        ranges.append((inst.texel_offset, inst.texel_offset + inst.texel_count))
        if side.is_double_sided:
            assert side.back_texel_offset == ranges[-1][1]
            ranges.append((side.back_texel_offset, side.back_texel_offset + inst.texel_count))
            expected_payload += 2 * inst.texel_count
        else:
            assert side.back_texel_offset == MESH_COLORS_INVALID_OFFSET
            expected_payload += inst.texel_count
        expected_surface += inst.texel_count
    for a, b in zip(ranges, ranges[1:]):
        assert a[1] == b[0]
    assert expected_surface == layout.total_surface_texels
    assert expected_payload == layout.total_payload_count


def test_material_flags_and_cornell_roles():
    # test_mesh_colors.py:84-105
    assert Material().flags == 0
    ds = Material(double_sided=True)
    assert ds.double_sided and ds.flags & MATERIAL_FLAG_DOUBLE_SIDED
    node = SceneNode.demo()
    for k, (_, mat, _) in enumerate(node.instances):
        assert node.materials[mat].double_sided == (k < 5)


def test_selective_back_payload_count_and_budget():
    # test_mesh_colors.py:107-136: 216 surface, 276 payload, 60 double-sided; 275 raises
    node = SceneNode.demo()
    layout = MeshColorsLayout.build(node, texels_per_unit=1.0, min_resolution=2, max_resolution=2,
                                    max_total_texels=276)
    assert layout.total_surface_texels == 216
    assert layout.total_payload_count == 276
    ds = sum(i.texel_count for i, s in zip(layout.instance_infos, layout.side_infos)
             if s.flags & MESH_COLORS_SIDE_FLAG_DOUBLE_SIDED)
    assert ds == 60
    with pytest.raises(ValueError, match="texel budget exceeded"):
        MeshColorsLayout.build(node, texels_per_unit=1.0, min_resolution=2, max_resolution=2, max_total_texels=275)
    with pytest.raises(ValueError, match="texel budget exceeded"):
        layout_oracle.build_layout(node, texels_per_unit=1.0, min_resolution=2, max_resolution=2,
                                   max_total_texels=275)


def test_budget_guard_rejects_impossible_layout():
    # test_mesh_colors.py:138-145
    with pytest.raises(ValueError, match="texel budget exceeded"):
        MeshColorsLayout.build(SceneNode.demo(), min_resolution=4, max_resolution=4, max_total_texels=1)


def test_resolution_limits_normalize_inward():
    # test_mesh_colors.py:147-163
    layout = MeshColorsLayout.build(SceneNode.demo(), min_resolution=3, max_resolution=7, max_total_texels=100_000)
    assert layout.min_resolution == 4 and layout.max_resolution == 4
    with pytest.raises(ValueError, match="min_resolution"):
        MeshColorsLayout.build(SceneNode.demo(), min_resolution=5, max_resolution=7, max_total_texels=100_000)


def test_cornell_config_totals():
    # 5 walls x 2 tris x 2145 (R=64) + light 2 x 153 (R=16) + tall block 8 x 561 + 4 x 153
    # + short block 12 x 153 = 28,692 surface; 21,450 back slots on the walls
    layout = MeshColorsLayout.build(cornell_box())
    walls = 5 * 2 * colors_per_patch(64)
    light = 2 * colors_per_patch(16)
    tall = 8 * colors_per_patch(32) + 4 * colors_per_patch(16)
    short = 12 * colors_per_patch(16)
    assert layout.total_surface_texels == walls + light + tall + short == 28_692
    assert layout.total_payload_count - layout.total_surface_texels == walls == 21_450
    assert layout.total_payload_count == 50_142


def _assert_same_layout(node, **kw):
    try:
        product = MeshColorsLayout.build(node, **kw)
    except ValueError as exc:
        with pytest.raises(ValueError, match=str(exc).split(";")[0]):
            layout_oracle.build_layout(node, **kw)
        return None
    f, i, s, ts, tp = layout_oracle.build_layout(node, **kw)
    np.testing.assert_array_equal(product.face_table, f)
    np.testing.assert_array_equal(product.instance_table, i)
    np.testing.assert_array_equal(product.side_table, s)
    assert (product.total_surface_texels, product.total_payload_count) == (ts, tp)
    return product


@pytest.mark.parametrize("kw", [
    {},
    dict(texels_per_unit=64.0, min_resolution=2, max_resolution=8, max_total_texels=100_000),
    dict(texels_per_unit=16.0, min_resolution=1, max_resolution=64, max_total_texels=30_000),
    dict(texels_per_unit=3.0, min_resolution=1, max_resolution=4, max_total_texels=300),
])
def test_product_layout_matches_oracle_cornell(kw):
    _assert_same_layout(cornell_box(), **kw)
    _assert_same_layout(SceneNode.demo(), **kw)


@pytest.mark.parametrize("seed,tris,budget", [(1, 3000, 40_000), (2, 8000, 20_000), (3, 2000, 1_000_000)])
def test_product_layout_matches_oracle_batch_scenes(seed, tris, budget):
    node = batch_scene(seed, tris)
    try:
        product = MeshColorsLayout.build(node, texels_per_unit=8.0, min_resolution=1, max_resolution=32,
                                         max_total_texels=budget)
    except ValueError as exc:
        with pytest.raises(ValueError):
            layout_oracle.build_layout(node, texels_per_unit=8.0, min_resolution=1, max_resolution=32,
                                       max_total_texels=budget)
        assert "budget" in str(exc)
        return
    f, i, s, ts, tp = layout_oracle.build_layout(node, texels_per_unit=8.0, min_resolution=1, max_resolution=32,
                                                 max_total_texels=budget)
    np.testing.assert_array_equal(product.face_table, f)
    np.testing.assert_array_equal(product.instance_table, i)
    np.testing.assert_array_equal(product.side_table, s)


def test_greedy_budget_halving_tail():
    # a budget that forces the last triangles to halve (the for/else path)
    node = cornell_box()
    full = MeshColorsLayout.build(node, max_total_texels=10_000_000)
    budget = full.total_payload_count - 100
    product = _assert_same_layout(node, max_total_texels=budget)
    assert product is not None and product.total_payload_count <= budget
    assert (product.face_table[:, 1] < full.face_table[:, 1]).any()


def test_texel_face_map_equals_binary_search():
    layout = MeshColorsLayout.build(cornell_box())
    tf = layout.texel_face_map()
    fi = layout.face_instance_map()
    prefix = layout.front_prefix()
    for w in range(0, layout.total_surface_texels, 97):
        f = int(tf[w])
        inst = int(fi[f])
        local = w - int(prefix[inst])
        a, r = layout.face_table[f]
        assert a <= local < a + colors_per_patch(int(r))


def test_synthetic_scene_sizes_and_layout():
    node = synthetic_scene(0)
    assert node.triangle_count == 524_288 + 49_152 + 640_000
    assert len(node.instances) == 1 + 4096 + 500
    layout = MeshColorsLayout.build(node, texels_per_unit=8.0, min_resolution=2, max_resolution=16,
                                    max_total_texels=16_777_216)
    assert layout.total_payload_count <= 16_777_216
    res = layout.face_table[:, 1]
    assert res.min() >= 2 and res.max() <= 16
