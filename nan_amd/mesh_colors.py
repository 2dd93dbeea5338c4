"""Mesh Colors layout and lattice math (mesh_colors.py:19-297 of the reference).

Every triangle carries a power-of-two barycentric lattice of
``(R+1)(R+2)/2`` irradiance texels.  ``MeshColorsLayout.build`` assigns
resolutions and a flat slot range per instance with the reference's greedy
budget rule; the tables it produces are bit-identical to the reference's
``uint2`` face, ``uint4`` instance and ``uint2`` side buffers.

The build is vectorised with numpy.  It reproduces the reference's
arithmetic: world positions are float32 ``M * (x, y, z, 1)`` products
(``spy.math.mul``), edge lengths are float64 norms of those float32
positions, and the budget rule is applied triangle by triangle in instance
order once the running total can exceed the budget.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .scene_node import SceneNode

MESH_COLORS_PAYLOAD_SIZE = 16
MESH_COLORS_INVALID_OFFSET = 0xFFFFFFFF
MESH_COLORS_SIDE_FLAG_DOUBLE_SIDED = 1 << 0

BUDGET_ERROR = (
    "Mesh Colors texel budget exceeded; lower --texture-space-texels-per-unit "
    "or --texture-space-min-resolution, or increase --texture-space-max-texels"
)


def colors_per_patch(resolution: int) -> int:
    """Lattice points of one triangle (mesh_colors.py:24-29)."""
    resolution = int(resolution)
    if resolution < 1:
        raise ValueError("resolution must be positive")
    return (resolution + 1) * (resolution + 2) // 2


def _row_offset(i: int, resolution: int) -> int:
    return i * (resolution + 1) - i * (i - 1) // 2


def address_to_ij(address: int, resolution: int) -> tuple[int, int]:
    """Inverse of :func:`ij_to_address` (mesh_colors.py:32-52).  Solved here
    by bisection on the row offsets, the form the shader uses
    (mesh_colors.slang:48-65); both forms agree on every address."""
    address = int(address)
    resolution = int(resolution)
    if address < 0 or address >= colors_per_patch(resolution):
        raise ValueError("address is outside the triangle patch")
    lo, hi = 0, resolution
    while lo < hi:
        mid = (lo + hi + 1) // 2
        if _row_offset(mid, resolution) <= address:
            lo = mid
        else:
            hi = mid - 1
    return lo, address - _row_offset(lo, resolution)


def ij_to_address(i: int, j: int, resolution: int) -> int:
    """Row-major lattice address (mesh_colors.py:55-61)."""
    i, j, resolution = int(i), int(j), int(resolution)
    if i < 0 or i > resolution or j < 0 or j > resolution - i:
        raise ValueError("(i, j) is outside the triangle patch")
    return _row_offset(i, resolution) + j


def barycentric_at(i: int, j: int, resolution: int) -> tuple[float, float, float]:
    """``(i/R, j/R, 1-i/R-j/R)`` (mesh_colors.py:64-69)."""
    if resolution <= 0:
        raise ValueError("resolution must be positive")
    u = float(i) / float(resolution)
    v = float(j) / float(resolution)
    return u, v, 1.0 - u - v


def next_power_of_two(value: int) -> int:
    value = max(1, int(value))
    return 1 << (value - 1).bit_length()


def previous_power_of_two(value: int) -> int:
    value = max(1, int(value))
    return 1 << (value.bit_length() - 1)


def _next_pow2_array(values: np.ndarray) -> np.ndarray:
    v = np.maximum(values.astype(np.int64), 1)
    out = np.ones_like(v)
    # Values here are bounded by max_resolution after clamping; a short loop
    # over bit positions is exact and fast.
    while True:
        mask = out < v
        if not mask.any():
            return out
        out[mask] <<= 1


@dataclass(frozen=True)
class MeshColorsFaceInfo:
    address: int
    resolution: int


@dataclass(frozen=True)
class MeshColorsInstanceInfo:
    face_offset: int
    face_count: int
    texel_offset: int
    texel_count: int


@dataclass(frozen=True)
class MeshColorsSideInfo:
    back_texel_offset: int
    flags: int

    @property
    def is_double_sided(self) -> bool:
        return bool(self.flags & MESH_COLORS_SIDE_FLAG_DOUBLE_SIDED)


def world_positions_f32(transform: np.ndarray, positions: np.ndarray) -> np.ndarray:
    """float32 ``M * (p, 1)`` with the dot products summed left to right."""
    m = np.asarray(transform, dtype=np.float32)
    p = np.asarray(positions, dtype=np.float32)
    out = np.empty(p.shape, dtype=np.float32)
    for r in range(3):
        acc = m[r, 0] * p[..., 0]
        acc = acc + m[r, 1] * p[..., 1]
        acc = acc + m[r, 2] * p[..., 2]
        acc = acc + m[r, 3]
        out[..., r] = acc
    return out


def instance_max_edges(scene_node: SceneNode) -> list[np.ndarray]:
    """Longest world-space edge of every triangle, per instance (float64 norms
    of the float32 ``M * (p, 1)`` positions, mesh_colors.py:215-224).
    Instances sharing a mesh are transformed together."""
    groups: dict[int, list[int]] = {}
    for k, (mesh_id, _, _) in enumerate(scene_node.instances):
        groups.setdefault(mesh_id, []).append(k)
    out: list[np.ndarray | None] = [None] * len(scene_node.instances)
    for mesh_id, members in groups.items():
        mesh = scene_node.meshes[mesh_id]
        p = mesh.vertices[:, 0:3].astype(np.float32)
        idx = mesh.indices.astype(np.int64)
        for start in range(0, len(members), max(1, (1 << 22) // max(1, len(p)))):
            chunk = members[start:start + max(1, (1 << 22) // max(1, len(p)))]
            m = np.stack([np.asarray(scene_node.transforms[scene_node.instances[k][2]].matrix, dtype=np.float32)
                          for k in chunk])
            world = np.empty((len(chunk), len(p), 3), dtype=np.float32)
            for r in range(3):
                acc = m[:, r, 0, None] * p[None, :, 0]
                acc = acc + m[:, r, 1, None] * p[None, :, 1]
                acc = acc + m[:, r, 2, None] * p[None, :, 2]
                acc = acc + m[:, r, 3, None]
                world[:, :, r] = acc
            w = world.astype(np.float64)
            a, b, c = w[:, idx[:, 0]], w[:, idx[:, 1]], w[:, idx[:, 2]]
            e0 = np.sqrt(np.sum((b - a) ** 2, axis=2))
            e1 = np.sqrt(np.sum((c - b) ** 2, axis=2))
            e2 = np.sqrt(np.sum((a - c) ** 2, axis=2))
            me = np.maximum(np.maximum(e0, e1), e2)
            for row, k in enumerate(chunk):
                out[k] = me[row]
    return out  # type: ignore[return-value]


def _resolutions(max_edge: np.ndarray, tpu: float, min_res: int, max_res: int) -> np.ndarray:
    """``min(max_res, nextpow2(max(min_res, ceil(max_edge * tpu))))``
    (mesh_colors.py:225-228)."""
    cells = np.ceil(max_edge * tpu)
    cells = np.maximum(cells, float(min_res))
    cells = np.minimum(cells, float(max_res))
    return np.minimum(_next_pow2_array(np.maximum(cells, 1).astype(np.int64)), max_res).astype(np.int64)


class MeshColorsLayout:
    """Face, instance and side tables of one scene (mesh_colors.py:121-297).

    The table arrays are ``face_table`` ``uint32[F,2]`` (address,
    resolution), ``instance_table`` ``uint32[I,4]`` (face_offset,
    face_count, texel_offset, texel_count) and ``side_table``
    ``uint32[I,2]`` (back_texel_offset, flags).
    """

    def __init__(self, face_table, instance_table, side_table, total_surface_texels,
                 total_payload_count, texels_per_unit, min_resolution, max_resolution,
                 adjacency_table=None, adjacency_diagnostics=None):
        from .mesh_colors_adjacency import MESH_COLORS_INVALID_ADJACENCY, MeshColorsAdjacencyDiagnostics

        self.face_table = np.ascontiguousarray(face_table, dtype=np.uint32)
        if adjacency_table is None:
            adjacency_table = np.full((len(self.face_table), 4), MESH_COLORS_INVALID_ADJACENCY, dtype=np.uint32)
        # uint32[F, 4]: packed edge records per global face, adjacent faces instance-local
        self.adjacency_table = np.ascontiguousarray(adjacency_table, dtype=np.uint32)
        self.adjacency_diagnostics = adjacency_diagnostics or MeshColorsAdjacencyDiagnostics()
        self.instance_table = np.ascontiguousarray(instance_table, dtype=np.uint32)
        self.side_table = np.ascontiguousarray(side_table, dtype=np.uint32)
        self.total_surface_texels = int(total_surface_texels)
        self.total_payload_count = int(total_payload_count)
        self.texels_per_unit = float(texels_per_unit)
        self.min_resolution = int(min_resolution)
        self.max_resolution = int(max_resolution)

    @property
    def total_texel_count(self) -> int:
        return self.total_surface_texels

    @property
    def face_infos(self) -> tuple[MeshColorsFaceInfo, ...]:
        return tuple(MeshColorsFaceInfo(int(a), int(r)) for a, r in self.face_table)

    @property
    def adjacency_infos(self):
        """Per global face, its three edge records (mesh_colors.py:124)."""
        from .mesh_colors_adjacency import _FaceView

        return _FaceView(self.adjacency_table)

    @property
    def instance_infos(self) -> tuple[MeshColorsInstanceInfo, ...]:
        return tuple(MeshColorsInstanceInfo(*(int(x) for x in row)) for row in self.instance_table)

    @property
    def side_infos(self) -> tuple[MeshColorsSideInfo, ...]:
        return tuple(MeshColorsSideInfo(int(a), int(f)) for a, f in self.side_table)

    def front_prefix(self) -> np.ndarray:
        """Exclusive prefix of front texel counts per instance (``uint32[I]``):
        the flat front-texel index of each instance's first texel."""
        counts = self.instance_table[:, 3].astype(np.int64)
        out = np.zeros(len(counts), dtype=np.int64)
        if len(counts) > 1:
            np.cumsum(counts[:-1], out=out[1:])
        return out.astype(np.uint32)

    def texel_face_map(self) -> np.ndarray:
        """Global face index of every front texel, in flat front-texel order
        (``uint32[total_surface_texels]``).  Faces tile each instance's front
        range contiguously, so this equals the result of the shader's
        per-texel binary search (mesh_colors.slang:272-305)."""
        res = self.face_table[:, 1].astype(np.int64)
        counts = (res + 1) * (res + 2) // 2
        return np.repeat(np.arange(len(res), dtype=np.uint32), counts)

    def face_instance_map(self) -> np.ndarray:
        """Instance index of every face (``uint32[F]``)."""
        return np.repeat(
            np.arange(len(self.instance_table), dtype=np.uint32),
            self.instance_table[:, 1].astype(np.int64),
        )

    @classmethod
    def build(cls, scene_node: SceneNode, *, texels_per_unit: float = 16.0, min_resolution: int = 4,
              max_resolution: int = 64, max_total_texels: int = 16_777_216) -> "MeshColorsLayout":
        """Greedy per-triangle resolution assignment (mesh_colors.py:147-297)."""
        texels_per_unit = max(float(texels_per_unit), 1e-6)
        min_resolution = next_power_of_two(max(1, int(min_resolution)))
        max_resolution = previous_power_of_two(max(1, int(max_resolution)))
        if min_resolution > max_resolution:
            raise ValueError(
                "Mesh Colors min_resolution must not exceed max_resolution "
                "after power-of-two normalization"
            )
        max_total_texels = max(1, int(max_total_texels))

        face_chunks: list[np.ndarray] = []
        instance_rows = []
        side_rows = []
        total_surface = 0
        total_payload = 0
        face_count_so_far = 0
        max_edges = instance_max_edges(scene_node)

        for mesh_id, material_id, transform_id in scene_node.instances:
            mesh = scene_node.meshes[mesh_id]
            material = scene_node.materials[material_id]
            double_sided = bool(getattr(material, "double_sided", False))
            mult = 2 if double_sided else 1
            res = _resolutions(max_edges[len(instance_rows)], texels_per_unit, min_resolution, max_resolution)
            patch = (res + 1) * (res + 2) // 2
            local_end = np.cumsum(patch)  # local address after each triangle
            over = total_payload + local_end * mult > max_total_texels
            if over.any():
                first = int(np.argmax(over))
                local = int(local_end[first - 1]) if first > 0 else 0
                for t in range(first, len(res)):
                    r = int(res[t])
                    p = (r + 1) * (r + 2) // 2
                    if total_payload + (local + p) * mult > max_total_texels:
                        candidate = r
                        while True:
                            if candidate <= min_resolution:
                                raise ValueError(BUDGET_ERROR)
                            candidate //= 2
                            p = (candidate + 1) * (candidate + 2) // 2
                            if total_payload + (local + p) * mult <= max_total_texels:
                                r = candidate
                                break
                        res[t] = r
                    local += p
                patch = (res + 1) * (res + 2) // 2
                local_end = np.cumsum(patch)
            n = len(res)
            texel_count = int(local_end[-1]) if n else 0
            addresses = np.zeros(n, dtype=np.int64)
            if n > 1:
                addresses[1:] = local_end[:-1]
            face_chunks.append(np.stack([addresses, res], axis=1))
            instance_rows.append((face_count_so_far, n, total_payload, texel_count))
            back = total_payload + texel_count if double_sided else MESH_COLORS_INVALID_OFFSET
            side_rows.append((back, MESH_COLORS_SIDE_FLAG_DOUBLE_SIDED if double_sided else 0))
            face_count_so_far += n
            total_surface += texel_count
            total_payload += texel_count * mult

        if total_surface == 0:
            raise ValueError("Texture-space rendering requires at least one triangle")
        if total_payload >= 1 << 32:
            raise ValueError("Mesh Colors payload exceeds 32-bit slot indices")

        faces = np.concatenate(face_chunks, axis=0) if face_chunks else np.zeros((0, 2))
        # edge adjacency of every mesh once, repeated per instance
        # (mesh_colors.py:174-206, 256)
        from .mesh_colors_adjacency import MeshColorsAdjacencyDiagnostics, build_triangle_adjacency

        mesh_adjacency = {}
        for mesh_id, _, _ in scene_node.instances:
            if mesh_id not in mesh_adjacency:
                mesh_adjacency[mesh_id] = build_triangle_adjacency(
                    np.asarray(scene_node.meshes[mesh_id].indices).reshape(-1, 3))
        diagnostics = MeshColorsAdjacencyDiagnostics()
        for adj in mesh_adjacency.values():
            diagnostics = diagnostics + adj.diagnostics
        adjacency = (np.concatenate([mesh_adjacency[m].packed for m, _, _ in scene_node.instances], axis=0)
                     if scene_node.instances else np.zeros((0, 4), dtype=np.uint32))
        return cls(
            faces.astype(np.uint32),
            np.array(instance_rows, dtype=np.uint64).astype(np.uint32).reshape(-1, 4),
            np.array(side_rows, dtype=np.uint64).astype(np.uint32).reshape(-1, 2),
            total_surface, total_payload, texels_per_unit, min_resolution, max_resolution,
            adjacency, diagnostics,
        )

    @classmethod
    def natural_demand(cls, scene_node: SceneNode, *, texels_per_unit: float = 16.0, min_resolution: int = 4,
                       max_resolution: int = 64, max_edges: list | None = None) -> int:
        """Payload slots the layout needs when no triangle is halved for the
        budget (build succeeds without halving iff this fits the budget).
        ``max_edges`` (from :func:`instance_max_edges`) may be passed to
        evaluate many parameter sets on one scene."""
        texels_per_unit = max(float(texels_per_unit), 1e-6)
        min_resolution = next_power_of_two(max(1, int(min_resolution)))
        max_resolution = previous_power_of_two(max(1, int(max_resolution)))
        if max_edges is None:
            max_edges = instance_max_edges(scene_node)
        total = 0
        for (mesh_id, material_id, _), edges in zip(scene_node.instances, max_edges):
            res = _resolutions(edges, texels_per_unit, min_resolution, max_resolution)
            mult = 2 if getattr(scene_node.materials[material_id], "double_sided", False) else 1
            total += int(((res + 1) * (res + 2) // 2).sum()) * mult
        return total

    def summary(self) -> str:
        """The reference's one-line layout report
        (texture_space_path_tracing_renderer.py:92-100)."""
        back = self.total_payload_count - self.total_surface_texels
        return (
            f"[TextureSpace] {len(self.instance_table)} instances, {len(self.face_table)} faces, "
            f"{self.total_surface_texels:,} surface texels, {back:,} back texels "
            f"({self.total_payload_count * MESH_COLORS_PAYLOAD_SIZE / (1024 * 1024):.1f} MiB writable; "
            f"{self.total_payload_count * 4 / (1024 * 1024):.1f} MiB RGB9E5)"
        )
