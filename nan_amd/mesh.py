"""Triangle meshes in the reference's container format.

A mesh holds ``float32[V, 8]`` vertices (object-space position, normal, uv)
and ``uint32[T, 3]`` indices, exactly as mesh.py:6-16 of the reference.
The quad and cube generators reproduce the reference's vertex order,
normals and winding (mesh.py:64-188) because the Mesh Colors layout and the
lattice orientation depend on both.
"""

from __future__ import annotations

import numpy as np


class Mesh:
    """Vertices ``float32[V,8]`` and indices ``uint32[T,3]`` (mesh.py:6-16)."""

    def __init__(self, vertices: np.ndarray, indices: np.ndarray):
        vertices = np.ascontiguousarray(vertices)
        indices = np.ascontiguousarray(indices)
        if vertices.ndim != 2 or vertices.shape[1] != 8 or vertices.dtype != np.float32:
            raise ValueError("Mesh vertices must be float32[V, 8]")
        if indices.ndim != 2 or indices.shape[1] != 3 or indices.dtype != np.uint32:
            raise ValueError("Mesh indices must be uint32[T, 3]")
        if indices.size and int(indices.max()) >= vertices.shape[0]:
            raise ValueError("Mesh index out of range")
        self.vertices = vertices
        self.indices = indices

    @property
    def vertex_count(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def triangle_count(self) -> int:
        return int(self.indices.shape[0])

    @property
    def index_count(self) -> int:
        return self.triangle_count * 3

    @classmethod
    def create_triangle(cls, pos_a, pos_b, pos_c, normal_a, normal_b, normal_c) -> "Mesh":
        """One triangle with per-vertex normals (mesh.py:30-62)."""
        vertices = np.array(
            [
                [*pos_a, *normal_a, 0, 0],
                [*pos_b, *normal_b, 1, 0],
                [*pos_c, *normal_c, 0, 1],
            ],
            dtype=np.float32,
        )
        return cls(vertices, np.array([[0, 1, 2]], dtype=np.uint32))

    @classmethod
    def create_quad(cls, size=(1.0, 1.0)) -> "Mesh":
        """XZ quad with normal +Y, winding ``[2,1,0],[1,2,3]`` (mesh.py:64-82)."""
        vertices = np.array(
            [
                [-0.5, 0, -0.5, 0, 1, 0, 0, 0],
                [+0.5, 0, -0.5, 0, 1, 0, 1, 0],
                [-0.5, 0, +0.5, 0, 1, 0, 0, 1],
                [+0.5, 0, +0.5, 0, 1, 0, 1, 1],
            ],
            dtype=np.float32,
        )
        vertices[:, 0] *= np.float32(size[0])
        vertices[:, 2] *= np.float32(size[1])
        return cls(vertices, np.array([[2, 1, 0], [1, 2, 3]], dtype=np.uint32))

    @classmethod
    def create_quad_yz(cls, size=(1.0, 1.0), face_positive_x: bool = True) -> "Mesh":
        """YZ quad facing +X or -X (mesh.py:84-105)."""
        nx = 1.0 if face_positive_x else -1.0
        vertices = np.array(
            [
                [0, -0.5, -0.5, nx, 0, 0, 0, 0],
                [0, +0.5, -0.5, nx, 0, 0, 1, 0],
                [0, -0.5, +0.5, nx, 0, 0, 0, 1],
                [0, +0.5, +0.5, nx, 0, 0, 1, 1],
            ],
            dtype=np.float32,
        )
        vertices[:, 1] *= np.float32(size[0])
        vertices[:, 2] *= np.float32(size[1])
        if face_positive_x:
            indices = np.array([[0, 1, 2], [3, 2, 1]], dtype=np.uint32)
        else:
            indices = np.array([[2, 1, 0], [1, 2, 3]], dtype=np.uint32)
        return cls(vertices, indices)

    @classmethod
    def create_quad_xy(cls, size=(1.0, 1.0), face_positive_z: bool = True) -> "Mesh":
        """XY quad facing +Z or -Z (mesh.py:107-125)."""
        nz = 1.0 if face_positive_z else -1.0
        vertices = np.array(
            [
                [-0.5, -0.5, 0, 0, 0, nz, 0, 0],
                [+0.5, -0.5, 0, 0, 0, nz, 1, 0],
                [-0.5, +0.5, 0, 0, 0, nz, 0, 1],
                [+0.5, +0.5, 0, 0, 0, nz, 1, 1],
            ],
            dtype=np.float32,
        )
        vertices[:, 0] *= np.float32(size[0])
        vertices[:, 1] *= np.float32(size[1])
        if face_positive_z:
            indices = np.array([[0, 1, 2], [3, 2, 1]], dtype=np.uint32)
        else:
            indices = np.array([[2, 1, 0], [1, 2, 3]], dtype=np.uint32)
        return cls(vertices, indices)

    @classmethod
    def create_cube(cls, size=(1.0, 1.0, 1.0)) -> "Mesh":
        """24-vertex cube with per-face normals and the reference's winding
        (mesh.py:127-188)."""
        faces = [
            # (positions of 4 corners, normal) in the reference's order
            ([(-0.5, -0.5, -0.5), (-0.5, -0.5, +0.5), (+0.5, -0.5, +0.5), (+0.5, -0.5, -0.5)], (0, -1, 0)),
            ([(-0.5, +0.5, +0.5), (-0.5, +0.5, -0.5), (+0.5, +0.5, -0.5), (+0.5, +0.5, +0.5)], (0, +1, 0)),
            ([(-0.5, +0.5, -0.5), (-0.5, -0.5, -0.5), (+0.5, -0.5, -0.5), (+0.5, +0.5, -0.5)], (0, 0, -1)),
            ([(+0.5, +0.5, +0.5), (+0.5, -0.5, +0.5), (-0.5, -0.5, +0.5), (-0.5, +0.5, +0.5)], (0, 0, +1)),
            ([(-0.5, +0.5, +0.5), (-0.5, -0.5, +0.5), (-0.5, -0.5, -0.5), (-0.5, +0.5, -0.5)], (-1, 0, 0)),
            ([(+0.5, +0.5, -0.5), (+0.5, -0.5, -0.5), (+0.5, -0.5, +0.5), (+0.5, +0.5, +0.5)], (+1, 0, 0)),
        ]
        uvs = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
        rows = []
        for corners, normal in faces:
            for corner, uv in zip(corners, uvs):
                rows.append([*corner, *normal, *uv])
        vertices = np.array(rows, dtype=np.float32)
        vertices[:, 0] *= np.float32(size[0])
        vertices[:, 1] *= np.float32(size[1])
        vertices[:, 2] *= np.float32(size[2])
        indices = []
        for face in range(6):
            b = face * 4
            indices.append([b + 0, b + 2, b + 1])
            indices.append([b + 0, b + 3, b + 2])
        return cls(vertices, np.array(indices, dtype=np.uint32))
