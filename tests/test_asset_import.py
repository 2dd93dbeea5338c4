"""Asset import (``SceneNode.load_asset``, scene_node.py:476-647) and the
scene description (``SceneNode.load_json``, called at app.py:151).  CPU only.

Known answers: the reference's demo scene written out as glTF/GLB loads
back to the same meshes, transforms and materials, so its Mesh Colors
layout repeats the reference's known answers (216 surface texels, 276
slots, 60 double-sided texels at one texel per unit and resolution 2;
test_mesh_colors.py:107-136) and the oracle traces it to the same bits as
the procedural scene.  Hand-made OBJ/STL/PLY/glTF files pin the parsing
rules (fan split, negative indices, strips, sparse accessors, node
hierarchies, byte strides, merged STL vertices)."""

from __future__ import annotations

import json
import math

import numpy as np
import pytest

import oracle
from nan_amd.material import ALPHA_MODE_MASK, Material
from nan_amd.mesh import Mesh
from nan_amd.mesh_colors import MeshColorsLayout
from nan_amd.scene_arrays import flatten_scene
from nan_amd.scene_node import SceneNode
from nan_amd.transform import Transform

import asset_files as W


def _known_answer_layout(node, budget=276):
    return MeshColorsLayout.build(node, texels_per_unit=1.0, min_resolution=2, max_resolution=2,
                                  max_total_texels=budget)


def _same_layout(a, b):
    np.testing.assert_array_equal(a.face_table, b.face_table)
    np.testing.assert_array_equal(a.instance_table, b.instance_table)
    np.testing.assert_array_equal(a.side_table, b.side_table)


@pytest.mark.parametrize("form", ["gltf_embedded", "gltf_external", "glb"])
def test_demo_round_trip_keeps_meshes_instances_and_layout(tmp_path, form):
    demo = SceneNode.demo()
    doc, blob = W.gltf_document(demo)
    path = tmp_path / ("demo.glb" if form == "glb" else "demo.gltf")
    if form == "glb":
        W.write_glb(str(path), doc, blob)
    else:
        W.write_gltf(str(path), doc, blob, embed=(form == "gltf_embedded"))
    node = SceneNode.load_asset(str(path))
    assert len(node.meshes) == len(demo.meshes) and len(node.instances) == len(demo.instances)
    for a, b in zip(node.meshes, demo.meshes):
        np.testing.assert_array_equal(a.vertices[:, 0:6].view(np.uint32), b.vertices[:, 0:6].view(np.uint32))
        np.testing.assert_array_equal(a.indices, b.indices)
        # uvs reach the scene as the file stores them
        np.testing.assert_array_equal(a.vertices[:, 6:8].view(np.uint32), b.vertices[:, 6:8].view(np.uint32))
    for (m, k, t), (dm, dk, dt) in zip(node.instances, demo.instances):
        assert m == dm
        np.testing.assert_array_equal(node.transforms[t].matrix.view(np.uint32),
                                      demo.transforms[dt].matrix.view(np.uint32))
        a, b = node.materials[k], demo.materials[dk]
        assert a.base_color == pytest.approx(b.base_color, rel=1e-7)
        assert a.emissive == b.emissive and a.double_sided == b.double_sided
    layout = _known_answer_layout(node)
    assert layout.total_surface_texels == 216 and layout.total_payload_count == 276
    assert sum(int(r[3]) for r, s in zip(layout.instance_table, layout.side_table) if s[1] & 1) == 60
    with pytest.raises(ValueError, match="texel budget exceeded"):
        _known_answer_layout(node, budget=275)
    _same_layout(layout, _known_answer_layout(demo))
    # camera frames the scene: the box spans [-1, 1] x [0, 2] x [-1, 1]
    d = math.sqrt(4 + 4 + 4)
    np.testing.assert_allclose(node.camera.target, [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(node.camera.position, [0, 1 + 0.3 * d, 1.5 * d], atol=1e-6)


def test_imported_demo_traces_to_the_same_bits(tmp_path):
    demo = SceneNode.demo()
    doc, blob = W.gltf_document(demo)
    W.write_glb(str(tmp_path / "demo.glb"), doc, blob)
    node = SceneNode.load_asset(str(tmp_path / "demo.glb"))
    out = []
    for n in (demo, node):
        layout = MeshColorsLayout.build(n, texels_per_unit=4.0, min_resolution=1, max_resolution=8)
        payload = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        oracle.OracleScene(flatten_scene(n)).trace_texels(layout.face_table, layout.instance_table,
                                                           layout.side_table, payload, 0, 1, 3, True,
                                                           threads=oracle.cpu_count())
        out.append(payload)
    np.testing.assert_array_equal(out[0].view(np.uint32), out[1].view(np.uint32))
    assert out[0][:, :3].max() > 0.0


def test_instancing_sorted_names_scale_and_auto_center(tmp_path):
    demo = SceneNode.demo()
    # node names in reverse order: instances come back sorted by node name
    names = [f"z{7 - i}" for i in range(len(demo.instances))]
    doc, blob = W.gltf_document(demo, node_names=names)
    W.write_gltf(str(tmp_path / "d.gltf"), doc, blob)
    node = SceneNode.load_asset(str(tmp_path / "d.gltf"), scale=0.5, auto_center=True)
    want = [demo.instances[i] for i in reversed(range(len(demo.instances)))]
    assert [m for m, _, _ in node.instances] == [m for m, _, _ in want]
    G = np.diag([0.5, 0.5, 0.5, 1.0]) @ np.array([[1, 0, 0, 0], [0, 1, 0, -1.0], [0, 0, 1, 0], [0, 0, 0, 1]])
    for (_, _, t), (_, _, dt) in zip(node.instances, want):
        np.testing.assert_array_equal(node.transforms[t].matrix,
                                      (G @ demo.transforms[dt].matrix.astype(np.float64)).astype(np.float32))
    np.testing.assert_allclose(node.camera.target, [0, 0, 0], atol=1e-7)
    d = math.sqrt(3.0)
    np.testing.assert_allclose(node.camera.position, [0, 0.3 * d, 1.5 * d], atol=1e-6)


def test_obj_world_space_demo_known_answers(tmp_path):
    demo = SceneNode.demo()
    W.write_obj(str(tmp_path / "demo.obj"), demo, material_names=["white", "red", "green", "box", "light"])
    node = SceneNode.load_asset(str(tmp_path / "demo.obj"))
    # one geometry per material, sorted by name, identity instances
    assert sorted(m.name for m in node.materials) == [m.name for m in node.materials]
    assert [m.name for m in node.materials] == ["box", "green", "light", "red", "white"]
    assert node.triangle_count == 36
    light = node.materials[2]
    assert light.emissive == (15.0, 15.0, 15.0)  # Ke kept
    assert node.materials[3].base_color == (0.65, 0.05, 0.05)  # Kd at full precision
    for _, _, t in node.instances:
        np.testing.assert_array_equal(node.transforms[t].matrix, np.eye(4, dtype=np.float32))
    # every world triangle comes back bit for bit, in material-group order
    want = {}
    for tris, k in W.world_triangles(demo):
        want.setdefault(k, []).append(tris)
    order = {"white": 0, "red": 1, "green": 2, "box": 3, "light": 4}
    for m, k, _ in node.instances:
        mesh = node.meshes[m]
        got = mesh.vertices[:, 0:3][mesh.indices.astype(np.int64)]
        np.testing.assert_array_equal(got, np.concatenate(want[order[node.materials[k].name]]))
    # OBJ has no double-sided flag: the front slots alone
    layout = _known_answer_layout(node)
    assert layout.total_surface_texels == 216 and layout.total_payload_count == 216


def test_obj_polygons_negative_indices_and_normals(tmp_path):
    (tmp_path / "q.obj").write_text(
        "# quad and a triangle\n"
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
        "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
        "vn 0 0 1\n"
        "f 1/1/1 2/2/1 3/3/1 4/4/1\n"
        "v 2 0 0\nv 3 0 0\nv 2 1 0 \\\n\n"
        "f -3 -2 \\\n -1\n")
    node = SceneNode.load_asset(str(tmp_path / "q.obj"))
    assert len(node.meshes) == 1
    mesh = node.meshes[0]
    # the fan (0,1,2), (0,2,3), then the triangle on its own new vertices
    np.testing.assert_array_equal(mesh.indices, [[0, 1, 2], [0, 2, 3], [4, 5, 6]])
    np.testing.assert_array_equal(mesh.vertices[4:, 0:3], [[2, 0, 0], [3, 0, 0], [2, 1, 0]])
    # one group mixing corners with and without normals: normals are computed, all +z for a flat patch
    np.testing.assert_array_equal(mesh.vertices[:, 3:6], np.tile([0, 0, 1], (7, 1)).astype(np.float32))
    assert node.materials[0].base_color == (0.8, 0.8, 0.8)


def test_obj_errors(tmp_path):
    with pytest.raises(FileNotFoundError):
        SceneNode.load_asset(str(tmp_path / "missing.obj"))
    (tmp_path / "e.obj").write_text("v 0 0 0\n")
    with pytest.raises(ValueError, match="no valid geometry"):
        SceneNode.load_asset(str(tmp_path / "e.obj"))
    (tmp_path / "bad.obj").write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(ValueError, match="out of range"):
        SceneNode.load_asset(str(tmp_path / "bad.obj"))
    (tmp_path / "x.fbx").write_text("")
    with pytest.raises(ValueError, match="unsupported file type"):
        SceneNode.load_asset(str(tmp_path / "x.fbx"))


def _cube_triangles():
    cube = Mesh.create_cube((1.0, 2.0, 3.0))
    return cube.vertices[:, 0:3][cube.indices.astype(np.int64)]


@pytest.mark.parametrize("kind", ["binary", "ascii"])
def test_stl_merges_vertices_centres_and_scales(tmp_path, kind):
    tris = _cube_triangles() + np.float32(5.0)
    path = str(tmp_path / f"c_{kind}.stl")
    (W.write_stl_binary if kind == "binary" else W.write_stl_ascii)(path, tris)
    node = SceneNode.load_asset(path)
    mesh = node.meshes[0]
    assert mesh.triangle_count == 12 and mesh.vertex_count == 8
    np.testing.assert_array_equal(mesh.vertices[:, 0:3][mesh.indices.astype(np.int64)], tris)
    # a cube corner's normal is the angle-weighted mean of three faces: the diagonal
    n = mesh.vertices[:, 3:6]
    np.testing.assert_allclose(np.abs(n), 1.0 / math.sqrt(3.0), rtol=1e-6)
    centred = SceneNode.load_asset(path, scale=2.0, auto_center=True).meshes[0]
    p = centred.vertices[:, 0:3]
    np.testing.assert_array_equal(p.min(axis=0), [-1.0, -2.0, -3.0])
    np.testing.assert_array_equal(p.max(axis=0), [1.0, 2.0, 3.0])


def test_stl_single_mesh_camera(tmp_path):
    W.write_stl_binary(str(tmp_path / "c.stl"), _cube_triangles())
    node = SceneNode.load_asset(str(tmp_path / "c.stl"))
    d = math.sqrt(1 + 4 + 9)
    np.testing.assert_allclose(node.camera.target, [0, 0, 0], atol=1e-7)
    np.testing.assert_allclose(node.camera.position, [0, 0.3 * d, 1.5 * d], rtol=1e-6)
    np.testing.assert_array_equal(node.transforms[0].matrix, np.eye(4, dtype=np.float32))


@pytest.mark.parametrize("binary", [False, True])
def test_ply_quads_and_file_normals(tmp_path, binary):
    pos = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1]]
    nrm = [[0, 0, 1], [0, 0, 1], [0, 0, 1], [0, 0, 1], [0.6, 0.8, 0]]
    faces = [[0, 1, 2, 3], [0, 1, 4]]
    path = str(tmp_path / "m.ply")
    W.write_ply(path, pos, nrm, faces, binary=binary)
    mesh = SceneNode.load_asset(path).meshes[0]
    np.testing.assert_array_equal(mesh.indices, [[0, 1, 2], [0, 2, 3], [0, 1, 4]])
    np.testing.assert_array_equal(mesh.vertices[:, 0:3], np.asarray(pos, dtype=np.float32))
    np.testing.assert_array_equal(mesh.vertices[:, 3:6], np.asarray(nrm, dtype=np.float32))


def _tri_doc(blob_parts, accessors, views, meshes, nodes, materials=None, scenes=True):
    blob = b"".join(blob_parts)
    doc = {"asset": {"version": "2.0"}, "accessors": accessors, "bufferViews": views,
           "buffers": [{"byteLength": len(blob)}], "meshes": meshes, "nodes": nodes}
    if materials:
        doc["materials"] = materials
    if scenes:
        doc["scenes"] = [{"nodes": [0]}]
        doc["scene"] = 0
    return doc, blob


def test_gltf_hierarchy_strip_sparse_stride_and_materials(tmp_path):
    # interleaved position+normal (stride 24) for 4 vertices, a uint8 strip
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], dtype=np.float32)
    nrm = np.tile(np.array([0, 0, 1], dtype=np.float32), (4, 1))
    inter = np.concatenate([pos, nrm], axis=1).tobytes()  # 96 B
    strip = np.array([0, 1, 2, 3], dtype=np.uint8).tobytes() + b"\0" * 4  # 8 B
    # sparse: vertex 3 moved to (2, 2, 0)
    sp_idx = np.array([3], dtype=np.uint16).tobytes() + b"\0\0"  # 4 B
    sp_val = np.array([2, 2, 0], dtype=np.float32).tobytes()  # 12 B
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": 96, "byteStride": 24},
             {"buffer": 0, "byteOffset": 96, "byteLength": 4},
             {"buffer": 0, "byteOffset": 104, "byteLength": 2},
             {"buffer": 0, "byteOffset": 108, "byteLength": 12}]
    accessors = [
        {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 0, "byteOffset": 12, "componentType": 5126, "count": 4, "type": "VEC3"},
        {"bufferView": 1, "componentType": 5121, "count": 4, "type": "SCALAR"},
        {"bufferView": 0, "componentType": 5126, "count": 4, "type": "VEC3",
         "sparse": {"count": 1, "indices": {"bufferView": 2, "componentType": 5123},
                    "values": {"bufferView": 3}}},
    ]
    meshes = [{"name": "s", "primitives": [
        {"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2, "mode": 5, "material": 0},
        {"attributes": {"POSITION": 3, "NORMAL": 1}, "indices": 2, "mode": 5, "material": 1}]}]
    q = (0.0, math.sin(math.pi / 4), 0.0, math.cos(math.pi / 4))  # 90 degrees about +y
    nodes = [{"name": "root", "translation": [10, 0, 0], "children": [1]},
             {"name": "child", "rotation": list(q), "scale": [2, 2, 2], "mesh": 0}]
    materials = [{"name": "lamp", "emissiveFactor": [1.0, 0.5, 0.25],
                  "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 8.0}},
                  "pbrMetallicRoughness": {"baseColorFactor": [0.2, 0.3, 0.4, 1.0], "metallicFactor": 0.0}},
                 {"name": "leaf", "alphaMode": "MASK", "alphaCutoff": 0.3, "doubleSided": True}]
    doc, blob = _tri_doc([inter, strip, sp_idx, sp_val], accessors, views, meshes, nodes, materials)
    W.write_glb(str(tmp_path / "h.glb"), doc, blob)
    node = SceneNode.load_asset(str(tmp_path / "h.glb"))
    assert len(node.meshes) == 2 and len(node.instances) == 2
    # strip (0,1,2), (2,1,3): every other triangle swaps its first two corners
    for mesh in node.meshes:
        np.testing.assert_array_equal(mesh.indices, [[0, 1, 2], [2, 1, 3]])
    np.testing.assert_array_equal(node.meshes[0].vertices[:, 0:3], pos)
    np.testing.assert_array_equal(node.meshes[1].vertices[3, 0:3], [2, 2, 0])
    np.testing.assert_array_equal(node.meshes[0].vertices[:, 3:6], nrm)
    # world = T(10,0,0) @ R_y(90) @ S(2): x -> -2 z, shifted by 10 in x
    want = np.array([[0, 0, 2, 10], [0, 2, 0, 0], [-2, 0, 0, 0], [0, 0, 0, 1]], dtype=np.float64)
    for _, _, t in node.instances:
        np.testing.assert_allclose(node.transforms[t].matrix, want, atol=1e-6)
    lamp, leaf = node.materials
    assert lamp.emissive == pytest.approx((8.0, 4.0, 2.0)) and lamp.base_color == pytest.approx((0.2, 0.3, 0.4))
    assert lamp.metallic == 0.0 and not lamp.double_sided
    assert leaf.double_sided and leaf.alpha_mode == ALPHA_MODE_MASK and leaf.alpha_cutoff == pytest.approx(0.3)
    assert leaf.base_color == (1.0, 1.0, 1.0)
    # the layout of a 2-instance scene: the double-sided leaf gets back slots
    layout = MeshColorsLayout.build(node, texels_per_unit=1.0, min_resolution=2, max_resolution=2)
    assert layout.total_surface_texels == 24 and layout.total_payload_count == 36


def test_gltf_without_indices_nodes_or_normals(tmp_path):
    pos = np.array([[0, 0, 0], [2, 0, 0], [0, 2, 0], [0, 0, 0], [0, 2, 0], [0, 0, 2]], dtype=np.float32)
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": pos.nbytes}]
    accessors = [{"bufferView": 0, "componentType": 5126, "count": 6, "type": "VEC3"}]
    meshes = [{"primitives": [{"attributes": {"POSITION": 0}}]}]
    nodes = [{"mesh": 0}]
    doc, blob = _tri_doc([pos.tobytes()], accessors, views, meshes, nodes, scenes=False)
    W.write_gltf(str(tmp_path / "n.gltf"), doc, blob, embed=False)
    mat = Material(base_color=(0.1, 0.2, 0.3))
    node = SceneNode.load_asset(str(tmp_path / "n.gltf"), default_material=mat)
    mesh = node.meshes[0]
    np.testing.assert_array_equal(mesh.indices, [[0, 1, 2], [3, 4, 5]])
    np.testing.assert_array_equal(mesh.vertices[:3, 3:6], np.tile([0, 0, 1], (3, 1)))
    np.testing.assert_array_equal(mesh.vertices[3:, 3:6], np.tile([1, 0, 0], (3, 1)))
    assert node.materials[0] is mat


def test_gltf_rejects_bad_indices(tmp_path):
    pos = np.zeros((3, 3), dtype=np.float32)
    idx = np.array([0, 1, 3], dtype=np.uint32)
    views = [{"buffer": 0, "byteOffset": 0, "byteLength": 36}, {"buffer": 0, "byteOffset": 36, "byteLength": 12}]
    accessors = [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                 {"bufferView": 1, "componentType": 5125, "count": 3, "type": "SCALAR"}]
    meshes = [{"primitives": [{"attributes": {"POSITION": 0}, "indices": 1}]}]
    doc, blob = _tri_doc([pos.tobytes(), idx.tobytes()], accessors, views, meshes, [{"mesh": 0}])
    W.write_glb(str(tmp_path / "b.glb"), doc, blob)
    with pytest.raises(ValueError, match="index out of range"):
        SceneNode.load_asset(str(tmp_path / "b.glb"))
    # an accessor that reads past its view
    accessors[1]["count"] = 4
    doc, blob = _tri_doc([pos.tobytes(), idx.tobytes()], accessors, views, meshes, [{"mesh": 0}])
    W.write_glb(str(tmp_path / "c.glb"), doc, blob)
    with pytest.raises(ValueError, match="past bufferView"):
        SceneNode.load_asset(str(tmp_path / "c.glb"))


def test_add_scene_node_and_load_json(tmp_path):
    demo = SceneNode.demo()
    doc, blob = W.gltf_document(demo)
    W.write_glb(str(tmp_path / "demo.glb"), doc, blob)
    W.write_stl_binary(str(tmp_path / "cube.stl"), _cube_triangles())
    desc = {"objects": [{"asset": "demo.glb"},
                        {"asset": "cube.stl", "translation": [5, 6, 7], "rotation": [0, 0, 90], "scale": 2}],
            "camera": {"position": [1, 2, 3], "target": [0, 0, 0], "fov": 45},
            "lighting": {"sun_direction": [0, 0, 1], "sun_intensity": 2.0, "atmosphere": True}}
    (tmp_path / "s.json").write_text(json.dumps(desc))
    node = SceneNode.load_json(str(tmp_path / "s.json"), axis_conversion="z_up_to_y_up")
    assert len(node.instances) == 9 and len(node.meshes) == 8
    # the demo keeps its own instance transforms (object transform identity)
    for (_, _, t), (_, _, dt) in zip(node.instances[:8], demo.instances):
        np.testing.assert_array_equal(node.transforms[t].matrix, demo.transforms[dt].matrix)
    m = node.transforms[node.instances[8][2]].matrix
    # z-up (5, 6, 7) -> y-up (5, 7, -6); rotation (0, 0, 90) -> (90, 0, 0) degrees about x
    np.testing.assert_allclose(m[:3, 3], [5, 7, -6])
    want = Transform((5, 7, -6), (math.radians(90), 0, 0), (2, 2, 2)).matrix
    np.testing.assert_array_equal(m, want)
    np.testing.assert_allclose(node.camera.position, [1, 3, -2])
    assert node.camera.fov == 45.0
    assert node.lighting.atmosphere and node.lighting.sun_direction == (0.0, 1.0, -0.0)
    # add_scene_node (the reference's behaviour): every added instance gets the one transform
    base = SceneNode.demo()
    ids = base.add_scene_node(SceneNode.load_asset(str(tmp_path / "cube.stl")), Transform((0, 3, 0)))
    assert ids == [8] and base.instances[8][0] == len(base.meshes) - 1
    ids = base.add_scene_node(SceneNode.demo(), Transform((0, 0, 5)))
    assert len(ids) == 8 and len({base.instances[i][2] for i in ids}) == 1
    with pytest.raises(ValueError, match="objects"):
        (tmp_path / "bad.json").write_text("{}")
        SceneNode.load_json(str(tmp_path / "bad.json"))


@pytest.mark.parametrize("where", ["bufferView", "data", "file"])
def test_gltf_textures_and_alpha_modes_are_imported(tmp_path, where):
    rng = np.random.default_rng(3)
    demo = SceneNode.demo()
    doc, blob = W.gltf_document(demo)
    base = rng.integers(0, 256, size=(6, 5, 4), dtype=np.uint8)
    normal = rng.integers(0, 256, size=(3, 4, 4), dtype=np.uint8)
    emit = rng.integers(0, 256, size=(2, 2, 4), dtype=np.uint8)
    blob = W.add_textures(doc, blob, {(0, "baseColorTexture"): base, (0, "normalTexture"): normal,
                                      (1, "emissiveTexture"): emit}, where=where, folder=str(tmp_path))
    doc["materials"][0]["alphaMode"] = "MASK"
    doc["materials"][0]["alphaCutoff"] = 0.25
    doc["materials"][1]["alphaMode"] = "BLEND"
    path = tmp_path / ("t.glb" if where == "bufferView" else "t.gltf")
    if where == "bufferView":
        W.write_glb(str(path), doc, blob)
    else:
        W.write_gltf(str(path), doc, blob, embed=True)
    node = SceneNode.load_asset(str(path))
    m0 = next(m for m in node.materials if m.name == "mat0")
    m1 = next(m for m in node.materials if m.name == "mat1")
    assert m0.alpha_mode == ALPHA_MODE_MASK and m0.alpha_cutoff == pytest.approx(0.25)
    assert m1.alpha_mode == 2 and m0.textured and m1.textured
    arrays = flatten_scene(node)
    tex = arrays["textures"]
    assert len(tex) == 3
    # the decoded pixels, R in the low byte, rows top to bottom as stored
    words = {}
    for first, w, h, flags in tex:
        words[(w, h)] = (arrays["texels"][first:first + w * h].reshape(h, w), flags)
    for img, srgb in ((base, 1), (normal, 0), (emit, 1)):
        got, flags = words[(img.shape[1], img.shape[0])]
        want = (img[..., 0].astype(np.uint32) | img[..., 1].astype(np.uint32) << 8 |
                img[..., 2].astype(np.uint32) << 16 | img[..., 3].astype(np.uint32) << 24)
        np.testing.assert_array_equal(got, want)
        assert flags == srgb
    mats = arrays["materials"]
    k0 = node.materials.index(m0)
    assert mats[k0, 7] == ALPHA_MODE_MASK and mats[k0, 8] != 0xFFFFFFFF and mats[k0, 10] != 0xFFFFFFFF
    assert mats[k0, 9] == 0xFFFFFFFF and mats[k0].view(np.float32)[11] == np.float32(0.25)


def test_missing_texture_file_leaves_the_slot_empty(tmp_path):
    demo = SceneNode.demo()
    doc, blob = W.gltf_document(demo)
    doc["images"] = [{"uri": "nowhere.png"}]
    doc["textures"] = [{"source": 0}]
    doc["materials"][0].setdefault("pbrMetallicRoughness", {})["baseColorTexture"] = {"index": 0}
    W.write_gltf(str(tmp_path / "m.gltf"), doc, blob, embed=True)
    node = SceneNode.load_asset(str(tmp_path / "m.gltf"))
    assert all(m.base_color_texture is None for m in node.materials)
    assert flatten_scene(node)["textures"].shape == (0, 4)
