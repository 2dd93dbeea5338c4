"""Flattening of a :class:`SceneNode` into the flat arrays the native
library consumes (the buffers scene.py:330-437 of the reference uploads).

* ``vertices``        float32[V,8]  all meshes concatenated (32-byte stride)
* ``indices``         uint32[T*3]   mesh-local indices, meshes concatenated
* ``mesh_descs``      uint32[M,4]   vertex_count, index_count, vertex_offset, index_offset
* ``instance_descs``  uint32[I,4]   mesh_id, material_id, transform_id, flags
* ``transforms``      float32[X,16] row-major object-to-world
* ``inv_transposes``  float32[X,16] row-major transpose(inverse(M))
* ``materials``       uint32[Mat,12] see ``Material.pack``
* ``lighting``        float32[16]   see ``Lighting.pack``
* ``textures``        uint32[Tx,4]  first texel, width, height, flags (``textures.pack_textures``)
* ``texels``          uint32[N]     RGBA8 texels of every texture
"""

from __future__ import annotations

from pathlib import Path

import numpy as np

from .material import MATERIAL_RECORD_WORDS, TEXTURE_SLOTS
from .textures import NO_TEXTURE, load_image, pack_textures
from .scene_node import SceneNode
from .transform import inverse_transpose, is_odd_negative

INSTANCE_FLAG_ODD_NEGATIVE_TRANSFORM = 1 << 0


def flatten_scene(scene_node: SceneNode) -> dict:
    if not scene_node.instances:
        raise ValueError("scene has no instances")
    mesh_descs = np.zeros((len(scene_node.meshes), 4), dtype=np.uint32)
    vcount = 0
    icount = 0
    for k, mesh in enumerate(scene_node.meshes):
        mesh_descs[k] = (mesh.vertex_count, mesh.index_count, vcount, icount)
        vcount += mesh.vertex_count
        icount += mesh.index_count
    vertices = np.concatenate([m.vertices for m in scene_node.meshes], axis=0).astype(np.float32)
    indices = np.concatenate([m.indices.reshape(-1) for m in scene_node.meshes]).astype(np.uint32)
    transforms = np.stack([np.asarray(t.matrix, dtype=np.float32).reshape(16) for t in scene_node.transforms])
    inv_t = np.stack([inverse_transpose(t.matrix).reshape(16) for t in scene_node.transforms])
    odd = np.array([is_odd_negative(t.matrix) for t in scene_node.transforms], dtype=bool)
    inst = np.asarray(scene_node.instances, dtype=np.uint32).reshape(-1, 3)
    instance_descs = np.zeros((len(inst), 4), dtype=np.uint32)
    instance_descs[:, 0:3] = inst
    instance_descs[:, 3] = np.where(odd[inst[:, 2]], INSTANCE_FLAG_ODD_NEGATIVE_TRANSFORM, 0)
    materials, textures, texels = _pack_materials(scene_node.materials)
    return {
        "vertices": np.ascontiguousarray(vertices),
        "indices": np.ascontiguousarray(indices),
        "mesh_descs": mesh_descs,
        "instance_descs": instance_descs,
        "transforms": np.ascontiguousarray(transforms, dtype=np.float32),
        "inv_transposes": np.ascontiguousarray(inv_t, dtype=np.float32),
        "materials": materials,
        "lighting": scene_node.lighting.pack(),
        "russian_roulette": bool(getattr(scene_node, "russian_roulette", False)),
        "textures": textures,
        "texels": texels,
    }


def _texture_key(src, srgb: bool):
    if isinstance(src, bytes):
        key = src
    elif isinstance(src, str) or hasattr(src, "__fspath__"):
        key = str(Path(src).resolve())
    else:
        key = id(src)  # an array: one texture per object
    return key, srgb


def _pack_materials(material_list) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Material records and the scene's textures.  A texture source used
    by several materials is loaded once per colour space (the reference's
    TextureManager caches by path, texture_manager.py:84-89)."""
    index: dict = {}
    images, srgb = [], []
    records = []
    for m in material_list:
        ids = []
        for name, is_srgb in TEXTURE_SLOTS:
            src = getattr(m, name, None)
            if src is None:
                ids.append(NO_TEXTURE)
                continue
            key = _texture_key(src, is_srgb)
            if key not in index:
                index[key] = len(images)
                images.append(load_image(src))
                srgb.append(is_srgb)
            ids.append(index[key])
        records.append(m.pack(tuple(ids)))
    if records:
        materials = np.stack(records).astype(np.uint32)
    else:
        materials = np.zeros((0, MATERIAL_RECORD_WORDS), dtype=np.uint32)
    texels, textures = pack_textures(images, srgb)
    return materials, textures, texels


def instance_triangle_offsets(arrays: dict) -> np.ndarray:
    """Global triangle id of each instance's first triangle (``uint32[I]``);
    global id = offset[instance] + primitive index."""
    tri_per_inst = arrays["mesh_descs"][arrays["instance_descs"][:, 0], 1].astype(np.int64) // 3
    out = np.zeros(len(tri_per_inst), dtype=np.int64)
    if len(out) > 1:
        np.cumsum(tri_per_inst[:-1], out=out[1:])
    return out.astype(np.uint32)


def world_bounds(arrays: dict) -> tuple[np.ndarray, np.ndarray]:
    """Conservative world-space bounding box (min, max float64[3]): each
    instance's object-space mesh box, its 8 corners transformed."""
    verts = arrays["vertices"][:, :3].astype(np.float64)
    descs = arrays["mesh_descs"].astype(np.int64)
    mesh_lo = np.empty((len(descs), 3))
    mesh_hi = np.empty((len(descs), 3))
    for k, (vc, _ic, vo, _io) in enumerate(descs):
        v = verts[vo:vo + vc]
        mesh_lo[k] = v.min(axis=0) if vc else 0.0
        mesh_hi[k] = v.max(axis=0) if vc else 0.0
    inst = arrays["instance_descs"].astype(np.int64)
    lo, hi = mesh_lo[inst[:, 0]], mesh_hi[inst[:, 0]]
    m = arrays["transforms"].astype(np.float64).reshape(-1, 4, 4)[inst[:, 2]]
    corners = np.stack([np.where(np.array([(c >> a) & 1 for a in range(3)], dtype=bool), hi, lo)
                        for c in range(8)], axis=1)  # [I, 8, 3]
    world = np.einsum("iab,icb->ica", m[:, :3, :3], corners) + m[:, None, :3, 3]
    return world.reshape(-1, 3).min(axis=0), world.reshape(-1, 3).max(axis=0)
