"""Layout oracle: a literal, loop-by-loop restatement of the reference's
``MeshColorsLayout.build`` (mesh_colors.py:147-297) and lattice helpers
(mesh_colors.py:24-79).

TEST INFRASTRUCTURE ONLY: the product never imports this module.

The structure follows the reference line by line: per instance, per
triangle, float32 ``M * (x, y, z, 1)`` world positions (the reference's
``spy.math.mul``), float64 edge norms, ``ceil``, power-of-two rounding,
and the greedy halving loop with its for/else error.  It is pinned by the
reference's known answers (test_mesh_colors.py:107-163) and by closed-form
totals in the tests.
"""

from __future__ import annotations

import math

import numpy as np

INVALID = 0xFFFFFFFF
SIDE_DOUBLE_SIDED = 1


def colors_per_patch(resolution: int) -> int:
    if resolution < 1:
        raise ValueError("resolution must be positive")
    return (resolution + 1) * (resolution + 2) // 2


def address_to_ij_sqrt(address: int, resolution: int) -> tuple[int, int]:
    """The reference's closed form with fix-up loops (mesh_colors.py:32-52)."""
    if address < 0 or address >= colors_per_patch(resolution):
        raise ValueError("address is outside the triangle patch")
    b = 2 * resolution + 3
    discriminant = b * b - 8 * address
    i = int((b - math.sqrt(discriminant)) * 0.5)
    cumulative = i * (resolution + 1) - i * (i - 1) // 2
    while cumulative > address:
        i -= 1
        cumulative = i * (resolution + 1) - i * (i - 1) // 2
    while i < resolution:
        next_cumulative = (i + 1) * (resolution + 1) - (i + 1) * i // 2
        if next_cumulative > address:
            break
        i += 1
        cumulative = next_cumulative
    return i, address - cumulative


def ij_to_address(i: int, j: int, resolution: int) -> int:
    if i < 0 or i > resolution or j < 0 or j > resolution - i:
        raise ValueError("(i, j) is outside the triangle patch")
    return i * (resolution + 1) - i * (i - 1) // 2 + j


def next_power_of_two(value: int) -> int:
    value = max(1, int(value))
    return 1 << (value - 1).bit_length()


def previous_power_of_two(value: int) -> int:
    value = max(1, int(value))
    return 1 << (value.bit_length() - 1)


def _world_position(m: np.ndarray, p) -> np.ndarray:
    f = np.float32
    out = []
    for r in range(3):
        acc = f(m[r, 0]) * f(p[0])
        acc = f(acc + f(m[r, 1]) * f(p[1]))
        acc = f(acc + f(m[r, 2]) * f(p[2]))
        acc = f(acc + f(m[r, 3]) * f(1.0))
        out.append(float(acc))
    return np.asarray(out, dtype=np.float64)


def build_layout(scene_node, *, texels_per_unit=16.0, min_resolution=4, max_resolution=64,
                 max_total_texels=16_777_216):
    """Returns (faces [F,2], instances [I,4], sides [I,2], total_surface, total_payload)."""
    texels_per_unit = max(float(texels_per_unit), 1e-6)
    min_resolution = next_power_of_two(max(1, int(min_resolution)))
    max_resolution = previous_power_of_two(max(1, int(max_resolution)))
    if min_resolution > max_resolution:
        raise ValueError("Mesh Colors min_resolution must not exceed max_resolution "
                         "after power-of-two normalization")
    max_total_texels = max(1, int(max_total_texels))
    faces, instances, sides = [], [], []
    total_surface = 0
    total_payload = 0
    for mesh_id, material_id, transform_id in scene_node.instances:
        mesh = scene_node.meshes[mesh_id]
        material = scene_node.materials[material_id]
        double_sided = bool(getattr(material, "double_sided", False))
        mult = 2 if double_sided else 1
        m = np.asarray(scene_node.transforms[transform_id].matrix, dtype=np.float32)
        face_offset = len(faces)
        texel_offset = total_payload
        local = 0
        for tri in mesh.indices:
            pos = [_world_position(m, mesh.vertices[int(k), 0:3]) for k in tri]
            max_edge = max(
                float(np.linalg.norm(pos[1] - pos[0])),
                float(np.linalg.norm(pos[2] - pos[1])),
                float(np.linalg.norm(pos[0] - pos[2])),
            )
            res = next_power_of_two(max(min_resolution, int(math.ceil(max_edge * texels_per_unit))))
            res = min(res, max_resolution)
            patch = colors_per_patch(res)
            if total_payload + (local + patch) * mult > max_total_texels:
                candidate = res
                while candidate > min_resolution:
                    candidate //= 2
                    patch = colors_per_patch(candidate)
                    if total_payload + (local + patch) * mult <= max_total_texels:
                        res = candidate
                        break
                else:
                    raise ValueError("Mesh Colors texel budget exceeded")
            faces.append((local, res))
            local += patch
        instances.append((face_offset, mesh.triangle_count, texel_offset, local))
        sides.append((texel_offset + local if double_sided else INVALID,
                      SIDE_DOUBLE_SIDED if double_sided else 0))
        total_surface += local
        total_payload += local * mult
    if total_surface == 0:
        raise ValueError("Texture-space rendering requires at least one triangle")
    return (np.asarray(faces, dtype=np.uint32).reshape(-1, 2),
            np.asarray(instances, dtype=np.uint32).reshape(-1, 4),
            np.asarray(sides, dtype=np.uint32).reshape(-1, 2),
            total_surface, total_payload)
