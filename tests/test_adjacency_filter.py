"""Triangle adjacency and the surface-filter oracle on the CPU, against the
reference's known answers (test_mesh_colors_adjacency.py,
test_mesh_colors_surface_filter.py:128-200) and invariants."""

from __future__ import annotations

import math

import numpy as np
import pytest

from nan_amd.material import Material
from nan_amd.mesh import Mesh
from nan_amd.mesh_colors import MeshColorsLayout, colors_per_patch
from nan_amd.mesh_colors_adjacency import (
    MESH_COLORS_ADJACENCY_FLIP_BIT,
    MESH_COLORS_INVALID_ADJACENCY,
    MeshColorsEdgeAdjacency,
    build_triangle_adjacency,
    edge_barycentrics,
    edge_lattice_ij,
    remap_edge_parameter,
    triangle_edge_vertices,
)
from nan_amd.mesh_colors_surface_filter import face_normals
from nan_amd.scene_node import SceneNode
from nan_amd.transform import Transform
from oracle import filter_oracle


def _valid_edges(adj) -> int:
    return sum(int(e.valid) for face in adj.faces for e in face.edges)


def _assert_reciprocal(indices, adj) -> None:
    tri = np.asarray(indices)
    for fi, face in enumerate(adj.faces):
        for ei, e in enumerate(face.edges):
            if not e.valid:
                continue
            rev = adj.faces[e.adjacent_face][e.adjacent_edge]
            assert rev.valid and rev.adjacent_face == fi and rev.adjacent_edge == ei and rev.flip == e.flip
            sv = triangle_edge_vertices(tri[fi], ei)
            av = triangle_edge_vertices(tri[e.adjacent_face], e.adjacent_edge)
            assert set(sv) == set(av)
            for t in (0.0, 0.25, 0.5, 1.0):
                at = remap_edge_parameter(t, e)
                assert {sv[0]: 1 - t, sv[1]: t} == pytest.approx({av[0]: 1 - at, av[1]: at})


# --- topology (test_mesh_colors_adjacency.py:69-210) -------------------------
def test_isolated_triangle_is_all_boundary():
    adj = build_triangle_adjacency(np.array([[0, 1, 2]], dtype=np.uint32))
    assert _valid_edges(adj) == 0
    assert adj.diagnostics.boundary_edge_count == 3 and adj.diagnostics.manifold_edge_count == 0


def test_two_triangle_quad_has_one_reciprocal_edge():
    idx = np.array([[0, 1, 2], [2, 1, 3]], dtype=np.uint32)
    adj = build_triangle_adjacency(idx)
    shared = adj.faces[0][1]
    assert (shared.adjacent_face, shared.adjacent_edge, shared.flip) == (1, 0, True)
    assert adj.diagnostics.boundary_edge_count == 4 and adj.diagnostics.manifold_edge_count == 1
    assert _valid_edges(adj) == 2
    _assert_reciprocal(idx, adj)


def test_strip_fans_and_tetrahedron():
    strip = np.array([[0, 1, 2], [2, 1, 3], [2, 3, 4]], dtype=np.uint32)
    adj = build_triangle_adjacency(strip)
    assert adj.faces[0][1].adjacent_face == 1 and adj.faces[1][2].adjacent_face == 2
    assert (adj.diagnostics.manifold_edge_count, adj.diagnostics.boundary_edge_count, _valid_edges(adj)) == (2, 5, 4)
    _assert_reciprocal(strip, adj)
    open_fan = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4]], dtype=np.uint32)
    adj = build_triangle_adjacency(open_fan)
    assert (adj.diagnostics.manifold_edge_count, adj.diagnostics.boundary_edge_count) == (2, 5)
    closed = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 1]], dtype=np.uint32)
    adj = build_triangle_adjacency(closed)
    assert (adj.diagnostics.manifold_edge_count, adj.diagnostics.boundary_edge_count) == (4, 4)
    face, walk = 0, []
    for _ in range(4):
        walk.append(face)
        face = adj.faces[face][2].adjacent_face
    assert walk == [0, 1, 2, 3] and face == 0
    _assert_reciprocal(closed, adj)
    tet = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], dtype=np.uint32)
    adj = build_triangle_adjacency(tet)
    d = adj.diagnostics
    assert (d.boundary_edge_count, d.manifold_edge_count, d.orientation_anomaly_count, _valid_edges(adj)) == (0, 6, 0, 12)
    _assert_reciprocal(tet, adj)


def test_winding_seams_non_manifold_and_degenerate():
    idx = np.array([[0, 1, 2], [1, 2, 3]], dtype=np.uint32)
    adj = build_triangle_adjacency(idx)
    assert adj.faces[0][1].valid and not adj.faces[0][1].flip
    assert adj.diagnostics.orientation_anomaly_count == 1
    _assert_reciprocal(idx, adj)
    adj = build_triangle_adjacency(np.array([[0, 1, 2], [3, 4, 5]], dtype=np.uint32))
    assert _valid_edges(adj) == 0 and adj.diagnostics.boundary_edge_count == 6
    adj = build_triangle_adjacency(np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4]], dtype=np.uint32))
    assert adj.diagnostics.non_manifold_edge_count == 1 and adj.diagnostics.boundary_edge_count == 6
    assert not any(adj.faces[k][0].valid for k in range(3))
    adj = build_triangle_adjacency(np.array([[0, 0, 1], [1, 2, 3]], dtype=np.uint32))
    assert adj.diagnostics.degenerate_face_count == 1 and adj.diagnostics.boundary_edge_count == 3
    assert _valid_edges(adj) == 0


def test_edge_coordinates_and_pack_round_trip():
    expected = {(0, 0.0): (1.0, 0.0, 0.0), (0, 1.0): (0.0, 1.0, 0.0), (1, 0.0): (0.0, 1.0, 0.0),
                (1, 1.0): (0.0, 0.0, 1.0), (2, 0.0): (0.0, 0.0, 1.0), (2, 1.0): (1.0, 0.0, 0.0)}
    for (edge, t), b in expected.items():
        assert edge_barycentrics(edge, t) == b
        assert edge_lattice_ij(edge, t, 8) == (b[0] * 8, b[1] * 8)
    for e in (MeshColorsEdgeAdjacency(), MeshColorsEdgeAdjacency(123, 2, False), MeshColorsEdgeAdjacency(456, 1, True)):
        assert MeshColorsEdgeAdjacency.unpack(e.pack()) == e
    # packed words as the kernel decodes them (test_mesh_colors_adjacency.py:290-352)
    packed = build_triangle_adjacency(np.array([[0, 1, 2], [2, 1, 3]], dtype=np.uint32)).packed_uint4()
    assert packed[0].tolist() == [MESH_COLORS_INVALID_ADJACENCY, (1 << 3) | 0 | MESH_COLORS_ADJACENCY_FLIP_BIT,
                                  MESH_COLORS_INVALID_ADJACENCY, MESH_COLORS_INVALID_ADJACENCY]


def test_random_meshes_match_a_dictionary_walk():
    """The sort-based pairing equals a literal walk of the reference's
    algorithm (edge uses grouped by sorted vertex pair, in face order)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        idx = rng.integers(0, 12, size=(int(rng.integers(1, 40)), 3))
        adj = build_triangle_adjacency(idx)
        uses = {}
        degenerate = 0
        for f, t in enumerate(idx.tolist()):
            if len(set(t)) != 3:
                degenerate += 1
                continue
            for e in range(3):
                a, b = t[e], t[(e + 1) % 3]
                uses.setdefault((min(a, b), max(a, b)), []).append((f, e, a, b))
        want = np.full((len(idx), 4), MESH_COLORS_INVALID_ADJACENCY, dtype=np.uint32)
        counts = [0, 0, 0, 0]
        for u in uses.values():
            if len(u) == 1:
                counts[0] += 1
                continue
            if len(u) != 2 or u[0][0] == u[1][0]:
                counts[2] += 1
                continue
            (f0, e0, s0, t0), (f1, e1, s1, t1) = u
            flip = s0 == t1 and t0 == s1
            counts[3] += 0 if flip else 1
            want[f0, e0] = MeshColorsEdgeAdjacency(f1, e1, flip).pack()
            want[f1, e1] = MeshColorsEdgeAdjacency(f0, e0, flip).pack()
            counts[1] += 1
        np.testing.assert_array_equal(adj.packed, want)
        d = adj.diagnostics
        assert (d.boundary_edge_count, d.manifold_edge_count, d.non_manifold_edge_count,
                d.orientation_anomaly_count, d.degenerate_face_count) == (*counts, degenerate)


def _node(positions, indices, double_sided=False, instances=1) -> SceneNode:
    positions = np.asarray(positions, dtype=np.float32)
    vertices = np.zeros((len(positions), 8), dtype=np.float32)
    vertices[:, :3] = positions
    node = SceneNode()
    mesh = node.add_mesh(Mesh(vertices, np.asarray(indices, dtype=np.uint32)))
    mat = node.add_material(Material(double_sided=double_sided))
    xf = node.add_transform(Transform())
    for _ in range(instances):
        node.add_instance(mesh, mat, xf)
    return node


def test_layout_repeats_mesh_local_adjacency_per_instance():
    node = _node([(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)], [[0, 1, 2], [2, 1, 3]], instances=2)
    layout = MeshColorsLayout.build(node, texels_per_unit=1.0, min_resolution=1, max_resolution=16,
                                    max_total_texels=10_000)
    assert len(layout.face_infos) == 4 and len(layout.adjacency_infos) == 4
    assert layout.adjacency_diagnostics.manifold_edge_count == 1
    assert layout.adjacency_diagnostics.boundary_edge_count == 4
    for inst in layout.instance_infos:
        n = layout.adjacency_infos[inst.face_offset][1]
        assert n.adjacent_face == 1
        assert inst.face_offset <= inst.face_offset + n.adjacent_face < inst.face_offset + inst.face_count


def test_adjacency_is_independent_of_neighbor_resolution():
    node = _node([(0, 0, 0), (1, 0, 0), (0, 1, 0), (8, 8, 0)], [[0, 1, 2], [1, 0, 3]])
    layout = MeshColorsLayout.build(node, texels_per_unit=1.0, min_resolution=1, max_resolution=16,
                                    max_total_texels=10_000)
    assert layout.face_infos[0].resolution == 2 and layout.face_infos[1].resolution == 16
    a = layout.adjacency_infos[0][0]
    assert (a.adjacent_face, a.adjacent_edge, a.flip) == (1, 0, True)
    assert edge_lattice_ij(0, 0.25, 2) == (1.5, 0.5)
    assert edge_lattice_ij(0, remap_edge_parameter(0.25, a), 16) == (4.0, 12.0)


# --- filter helpers (test_mesh_colors_surface_filter.py:128-180) -------------
def test_edge_tap_remap_known_answers():
    ok, b = filter_oracle.map_one_edge_tap(3, 2, 4, 0, True, 0)
    assert ok
    np.testing.assert_allclose(b, [0.25, 0.5, 0.25])
    ok, b = filter_oracle.map_one_edge_tap(2, 2, 3, 0, True, 0)
    assert ok
    np.testing.assert_allclose(b, [1 / 3, 1 / 3, 1 / 3], atol=1e-6)
    assert not np.allclose(b * 8.0, np.round(b * 8.0))  # off-lattice on the neighbour
    ok, _ = filter_oracle.map_one_edge_tap(-1, 0, 4, 0, True, 1)
    assert not ok


def test_normal_weight_formula():
    sigma = math.pi / 6.0
    z, y = (0.0, 0.0, 1.0), (0.0, 1.0, 0.0)
    d = np.array([0.0, 1.0, 1.0]) / math.sqrt(2.0)
    assert filter_oracle.normal_filter_weight(z, z, sigma) == pytest.approx(1.0, abs=1e-6)
    assert filter_oracle.normal_filter_weight(z, y, sigma) == pytest.approx(math.exp(-4.5), abs=1e-6)
    assert filter_oracle.normal_filter_weight(z, d, sigma) == pytest.approx(math.exp(-1.125), abs=1e-6)


def _oracle_pass(node, layout, payload, passes=1, sigma_spatial=1.0, sigma_deg=30.0):
    normals = face_normals(node, layout)
    nw = math.exp(-0.5 / sigma_spatial ** 2)
    src = payload
    for _ in range(passes):
        src = filter_oracle.surface_filter(layout.face_table, layout.instance_table, layout.side_table,
                                           layout.adjacency_table, normals, src, nw, math.radians(sigma_deg))
    return src


def _payload(layout, values, count=17):
    data = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    data[:, :3] = values
    data.view(np.uint32)[:, 3] = count
    return data


def test_oracle_preserves_constant_fields_and_isolates_seams():
    fan = _node([(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0)],
                [[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 1]])
    layout = MeshColorsLayout.build(fan, texels_per_unit=1.0, min_resolution=4, max_resolution=4,
                                    max_total_texels=100_000)
    out = _oracle_pass(fan, layout, _payload(layout, [2.0, 0.5, 4.0]), passes=4)
    np.testing.assert_allclose(out[:, :3], np.broadcast_to([2.0, 0.5, 4.0], out[:, :3].shape), atol=2e-6)
    assert (out.view(np.uint32)[:, 3] == 17).all()
    seam = _node([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 0), (1, 0, 0), (0, -1, 0)], [[0, 1, 2], [3, 4, 5]])
    layout = MeshColorsLayout.build(seam, texels_per_unit=1.0, min_resolution=2, max_resolution=2,
                                    max_total_texels=100_000)
    values = np.zeros((layout.total_payload_count, 3), dtype=np.float32)
    values[:colors_per_patch(2)] = 1.0
    out = _oracle_pass(seam, layout, _payload(layout, values), passes=3)
    assert out[colors_per_patch(2):, :3].max() == 0.0
