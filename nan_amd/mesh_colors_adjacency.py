"""Triangle edge adjacency for Mesh Colors filtering (mesh_colors_adjacency.py
of the reference).

Per face, three packed edge records (plus one unused word, so a face is one
``uint4``): ``adjacent_face << 3 | adjacent_edge | flip << 2`` with the
adjacent face mesh-local, or ``0xFFFFFFFF`` for a boundary, non-manifold or
degenerate edge.  Edge ``e`` of a triangle runs from vertex ``e`` to vertex
``(e+1) % 3``; two uses of an edge are paired in face order, and ``flip``
says the neighbour traverses it in the opposite direction (consistent
winding).  The pairing is computed with one sort over all edge uses instead
of a dictionary walk, and gives the same table and diagnostics.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MESH_COLORS_EDGE_COUNT = 3
MESH_COLORS_INVALID_ADJACENCY = 0xFFFFFFFF
MESH_COLORS_ADJACENCY_FACE_SHIFT = 3
MESH_COLORS_ADJACENCY_EDGE_MASK = 0x3
MESH_COLORS_ADJACENCY_FLIP_BIT = 1 << 2
MESH_COLORS_ADJACENCY_MAX_FACE = (1 << (32 - MESH_COLORS_ADJACENCY_FACE_SHIFT)) - 1


@dataclass(frozen=True)
class MeshColorsEdgeAdjacency:
    """One edge record (mesh_colors_adjacency.py:20-57)."""

    adjacent_face: int = MESH_COLORS_INVALID_ADJACENCY
    adjacent_edge: int = 0
    flip: bool = False

    @property
    def valid(self) -> bool:
        return self.adjacent_face != MESH_COLORS_INVALID_ADJACENCY

    def pack(self) -> int:
        if not self.valid:
            return MESH_COLORS_INVALID_ADJACENCY
        face, edge = int(self.adjacent_face), int(self.adjacent_edge)
        if face < 0 or face > MESH_COLORS_ADJACENCY_MAX_FACE:
            raise ValueError("adjacent face does not fit in the packed format")
        if edge < 0 or edge >= MESH_COLORS_EDGE_COUNT:
            raise ValueError("adjacent edge must be in [0, 2]")
        return face << MESH_COLORS_ADJACENCY_FACE_SHIFT | edge | (MESH_COLORS_ADJACENCY_FLIP_BIT if self.flip else 0)

    @classmethod
    def unpack(cls, packed: int) -> "MeshColorsEdgeAdjacency":
        packed = int(packed) & 0xFFFFFFFF
        if packed == MESH_COLORS_INVALID_ADJACENCY:
            return cls()
        edge = packed & MESH_COLORS_ADJACENCY_EDGE_MASK
        if edge >= MESH_COLORS_EDGE_COUNT:
            raise ValueError("packed adjacency contains an invalid edge")
        return cls(packed >> MESH_COLORS_ADJACENCY_FACE_SHIFT, edge, bool(packed & MESH_COLORS_ADJACENCY_FLIP_BIT))


@dataclass(frozen=True)
class MeshColorsFaceAdjacency:
    edges: tuple

    def __getitem__(self, edge: int) -> MeshColorsEdgeAdjacency:
        return self.edges[int(edge)]

    def pack(self) -> bytes:
        return np.array([e.pack() for e in self.edges] + [MESH_COLORS_INVALID_ADJACENCY], dtype=np.uint32).tobytes()


@dataclass(frozen=True)
class MeshColorsAdjacencyDiagnostics:
    boundary_edge_count: int = 0
    manifold_edge_count: int = 0
    non_manifold_edge_count: int = 0
    degenerate_face_count: int = 0
    orientation_anomaly_count: int = 0

    def __add__(self, other: "MeshColorsAdjacencyDiagnostics") -> "MeshColorsAdjacencyDiagnostics":
        return MeshColorsAdjacencyDiagnostics(*(a + b for a, b in zip(_fields(self), _fields(other))))


def _fields(d: MeshColorsAdjacencyDiagnostics) -> tuple:
    return (d.boundary_edge_count, d.manifold_edge_count, d.non_manifold_edge_count, d.degenerate_face_count,
            d.orientation_anomaly_count)


class _FaceView:
    """Indexable view of a packed table as MeshColorsFaceAdjacency records."""

    def __init__(self, packed: np.ndarray):
        self._packed = packed

    def __len__(self) -> int:
        return len(self._packed)

    def __getitem__(self, face: int) -> MeshColorsFaceAdjacency:
        row = self._packed[int(face)]
        return MeshColorsFaceAdjacency(tuple(MeshColorsEdgeAdjacency.unpack(int(row[e])) for e in range(3)))

    def __iter__(self):
        for k in range(len(self)):
            yield self[k]


class MeshColorsMeshAdjacency:
    """Adjacency of one mesh: ``packed`` uint32[F, 4] and diagnostics."""

    def __init__(self, packed: np.ndarray, diagnostics: MeshColorsAdjacencyDiagnostics):
        self.packed = packed
        self.diagnostics = diagnostics

    @property
    def faces(self) -> _FaceView:
        return _FaceView(self.packed)

    def packed_uint4(self) -> np.ndarray:
        return self.packed.copy()


def triangle_edge_vertices(triangle, edge: int) -> tuple[int, int]:
    vertices = np.asarray(triangle).reshape(-1)
    if vertices.size != 3:
        raise ValueError("triangle must contain exactly three vertex indices")
    edge = int(edge)
    if edge < 0 or edge >= MESH_COLORS_EDGE_COUNT:
        raise ValueError("edge must be in [0, 2]")
    return int(vertices[edge]), int(vertices[(edge + 1) % 3])


def edge_barycentrics(edge: int, t: float) -> tuple[float, float, float]:
    """Barycentrics on an edge, parameterised in vertex order."""
    edge, t = int(edge), float(t)
    if edge == 0:
        return 1.0 - t, t, 0.0
    if edge == 1:
        return 0.0, 1.0 - t, t
    if edge == 2:
        return t, 0.0, 1.0 - t
    raise ValueError("edge must be in [0, 2]")


def edge_lattice_ij(edge: int, t: float, resolution: int) -> tuple[float, float]:
    b = edge_barycentrics(edge, t)
    resolution = int(resolution)
    if resolution < 1:
        raise ValueError("resolution must be positive")
    return b[0] * resolution, b[1] * resolution


def remap_edge_parameter(t: float, adjacency: MeshColorsEdgeAdjacency) -> float:
    if not adjacency.valid:
        raise ValueError("cannot remap across a boundary edge")
    return 1.0 - float(t) if adjacency.flip else float(t)


def build_triangle_adjacency(indices) -> MeshColorsMeshAdjacency:
    """Edge adjacency of an indexed triangle mesh
    (mesh_colors_adjacency.py:164-246)."""
    tri = np.asarray(indices)
    if tri.ndim != 2 or tri.shape[1] != 3:
        raise ValueError("indices must have shape (triangle_count, 3)")
    if not np.issubdtype(tri.dtype, np.integer):
        raise ValueError("indices must use an integer dtype")
    tri = tri.astype(np.int64)
    f = len(tri)
    packed = np.full((f, 4), MESH_COLORS_INVALID_ADJACENCY, dtype=np.uint32)
    degenerate = (tri[:, 0] == tri[:, 1]) | (tri[:, 1] == tri[:, 2]) | (tri[:, 0] == tri[:, 2])
    faces = np.nonzero(~degenerate)[0]
    if len(faces) == 0:
        return MeshColorsMeshAdjacency(packed, MeshColorsAdjacencyDiagnostics(degenerate_face_count=int(degenerate.sum())))
    # every use of an edge, in the reference's iteration order (face, then edge)
    face = np.repeat(faces, 3)
    edge = np.tile(np.arange(3, dtype=np.int64), len(faces))
    start = tri[face, edge]
    end = tri[face, (edge + 1) % 3]
    lo, hi = np.minimum(start, end), np.maximum(start, end)
    order = np.lexsort((face * 3 + edge, hi, lo))
    lo, hi, face, edge, start, end = lo[order], hi[order], face[order], edge[order], start[order], end[order]
    new_group = np.ones(len(lo), dtype=bool)
    new_group[1:] = (lo[1:] != lo[:-1]) | (hi[1:] != hi[:-1])
    group_start = np.nonzero(new_group)[0]
    group_size = np.diff(np.append(group_start, len(lo)))
    pair = group_size == 2
    first = group_start[pair]
    second = first + 1
    manifold = face[first] != face[second]
    first, second = first[manifold], second[manifold]
    flip = (start[first] == end[second]) & (end[first] == start[second])
    if len(first) and max(int(face[first].max()), int(face[second].max())) > MESH_COLORS_ADJACENCY_MAX_FACE:
        raise ValueError("adjacent face does not fit in the packed format")
    flip_bits = np.where(flip, MESH_COLORS_ADJACENCY_FLIP_BIT, 0).astype(np.int64)
    packed[face[first], edge[first]] = ((face[second] << MESH_COLORS_ADJACENCY_FACE_SHIFT) | edge[second]
                                        | flip_bits).astype(np.uint32)
    packed[face[second], edge[second]] = ((face[first] << MESH_COLORS_ADJACENCY_FACE_SHIFT) | edge[first]
                                          | flip_bits).astype(np.uint32)
    diagnostics = MeshColorsAdjacencyDiagnostics(
        boundary_edge_count=int((group_size == 1).sum()),
        manifold_edge_count=int(len(first)),
        non_manifold_edge_count=int(len(group_size) - (group_size == 1).sum() - len(first)),
        degenerate_face_count=int(degenerate.sum()),
        orientation_anomaly_count=int((~flip).sum()),
    )
    return MeshColorsMeshAdjacency(packed, diagnostics)
