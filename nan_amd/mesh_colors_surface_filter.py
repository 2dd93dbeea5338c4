"""Mesh Colors surface filter pass (mesh_colors_surface_filter.py of the
reference): one-ring, normal-aware diffusion of the texel cache before the
RGB9E5 freeze, run by the gfx950 kernel behind ``nan_surface_filter``.

The reference dispatches the kernel once per instance and finds each
texel's face by binary search; here one launch covers every texel through
the flat texel->face map a native tracer object holds for the layout.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _native
from .device import CommandEncoder, Device
from .mesh_colors import MeshColorsLayout, world_positions_f32
from .texture_space_path_tracer import TextureSpacePathTracer, as_payload_tensor


def face_normals(scene_node, layout: MeshColorsLayout) -> np.ndarray:
    """World-space unit normal of every face, float32[F, 4]
    (mesh_colors_surface_filter.py:53-87): float32 ``M * (p, 1)`` vertex
    positions, cross product and normalisation in float64, zero for
    degenerate faces."""
    out = []
    for mesh_id, _, transform_id in scene_node.instances:
        mesh = scene_node.meshes[mesh_id]
        tri = np.asarray(mesh.indices, dtype=np.int64).reshape(-1, 3)
        world = world_positions_f32(np.asarray(scene_node.transforms[transform_id].matrix),
                                    np.asarray(mesh.vertices)[:, :3]).astype(np.float64)
        p = world[tri]
        n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
        length = np.linalg.norm(n, axis=1, keepdims=True)
        n = np.where(length > 1e-20, n / np.where(length > 1e-20, length, 1.0), 0.0)
        out.append(np.concatenate([n, np.zeros((len(n), 1))], axis=1))
    normals = np.concatenate(out, axis=0).astype(np.float32) if out else np.zeros((0, 4), np.float32)
    if len(normals) != len(layout.face_table):
        raise ValueError("Mesh Colors face-normal count does not match layout")
    return np.ascontiguousarray(normals)


class MeshColorsSurfaceFilter:
    """One-ring, normal-aware Mesh Colors diffusion pass
    (mesh_colors_surface_filter.py:12-128).  ``texture_space_path_tracer``
    (an addition) shares the tracer's native layout tables; without it the
    pass builds its own from ``scene`` (a :class:`~nan_amd.scene.Scene`, or
    any object with a ``scene_node``)."""

    def __init__(self, device: Device, scene, layout: MeshColorsLayout, *, face_infos_buffer=None,
                 instance_infos_buffer=None, side_infos_buffer=None, adjacency_infos_buffer=None,
                 face_normals_buffer=None, texture_space_path_tracer: TextureSpacePathTracer | None = None):
        self.device = device
        self.layout = layout
        dev = device.torch_device
        if texture_space_path_tracer is None:
            from .scene import Scene

            native_scene = scene if hasattr(scene, "handle") else Scene(device, scene.scene_node)
            texture_space_path_tracer = TextureSpacePathTracer(device, native_scene, layout)
        self.texture_space_path_tracer = texture_space_path_tracer
        self.face_infos_buffer = face_infos_buffer
        self.instance_infos_buffer = instance_infos_buffer
        self.side_infos_buffer = side_infos_buffer
        if adjacency_infos_buffer is None:
            adjacency_infos_buffer = torch.from_numpy(layout.adjacency_table.view(np.int32)).to(dev)
        self.adjacency_infos_buffer = adjacency_infos_buffer
        if face_normals_buffer is None:
            face_normals_buffer = torch.from_numpy(face_normals(scene.scene_node, layout)).to(dev)
        self.face_normals_buffer = face_normals_buffer

    def execute(self, command_encoder: CommandEncoder, source: torch.Tensor, destination: torch.Tensor, *,
                spatial_sigma: float = 1.0, normal_sigma_radians: float = math.radians(30.0)) -> None:
        if source is destination or source.data_ptr() == destination.data_ptr():
            raise ValueError("Mesh Colors surface filtering requires ping-pong buffers")
        as_payload_tensor(source, self.layout.total_payload_count)
        as_payload_tensor(destination, self.layout.total_payload_count)
        spatial_sigma = max(float(spatial_sigma), 1e-6)
        neighbor_weight = math.exp(-0.5 / (spatial_sigma * spatial_sigma))
        normal_sigma_radians = max(float(normal_sigma_radians), 0.0)
        _native.check(_native.lib().nan_surface_filter(
            self.texture_space_path_tracer.handle, self.adjacency_infos_buffer.data_ptr(),
            self.face_normals_buffer.data_ptr(), source.data_ptr(), destination.data_ptr(), neighbor_weight,
            normal_sigma_radians, command_encoder.stream_handle), "nan_surface_filter")
