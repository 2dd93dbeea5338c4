"""Affine transforms (transform.py:1-17 of the reference).

``matrix = T * R * S`` with ``R = Rz * Ry * Rx`` (angles in radians, X
applied first).  The matrix is assembled in float64 and rounded once to
float32; entries whose magnitude is below 1e-12 (the residue of sin(pi) and
the like) are snapped to zero so that axis-aligned scenes stay exactly
axis-aligned.
"""

from __future__ import annotations

import math

import numpy as np


def _rot_x(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]], dtype=np.float64)


def _rot_y(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1]], dtype=np.float64)


def _rot_z(a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)


class Transform:
    """Translation, rotation (radians) and scaling composed as ``T*R*S``."""

    def __init__(self, translation=(0.0, 0.0, 0.0), rotation=(0.0, 0.0, 0.0), scaling=(1.0, 1.0, 1.0)):
        self.translation = tuple(float(v) for v in translation)
        self.rotation = tuple(float(v) for v in rotation)
        self.scaling = tuple(float(v) for v in scaling)
        self.matrix = np.eye(4, dtype=np.float32)
        self.update_matrix()

    def update_matrix(self) -> None:
        t = np.eye(4, dtype=np.float64)
        t[0:3, 3] = self.translation
        s = np.diag([*self.scaling, 1.0])
        r = _rot_z(self.rotation[2]) @ _rot_y(self.rotation[1]) @ _rot_x(self.rotation[0])
        m = t @ r @ s
        m[np.abs(m) < 1e-12] = 0.0
        self.matrix = m.astype(np.float32)

    @classmethod
    def from_matrix(cls, matrix) -> "Transform":
        out = cls()
        out.matrix = np.asarray(matrix, dtype=np.float32).reshape(4, 4).copy()
        return out


def inverse_transpose(matrix: np.ndarray) -> np.ndarray:
    """``transpose(inverse(M))`` computed in float64, rounded to float32
    (scene.py:426-428 computes the same product in float32)."""
    inv = np.linalg.inv(np.asarray(matrix, dtype=np.float64))
    out = inv.T.copy()
    out[np.abs(out) < 1e-12] = 0.0
    return out.astype(np.float32)


def is_odd_negative(matrix: np.ndarray) -> bool:
    """Determinant sign test of the upper 3x3 (scene.py:350-371): ``<= 0``
    sets the odd-negative flag."""
    m = np.asarray(matrix, dtype=np.float64)
    return float(np.dot(np.cross(m[0, 0:3], m[1, 0:3]), m[2, 0:3])) <= 0.0
