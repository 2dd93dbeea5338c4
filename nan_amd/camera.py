"""Pinhole camera with Halton(2,3) sub-pixel jitter (camera.py:1-117 of the
reference).

Only what the primary-ray resolve consumes is kept: position, the image
plane vectors ``image_u/v/w``, jitter, and the near/far clip planes
(camera.slang:13-66).  ``pack()`` produces the 16-float record the native
resolve kernel reads.
"""

from __future__ import annotations

import math

import numpy as np


def halton(index: int, base: int) -> float:
    """Radical inverse of ``index`` in ``base`` (camera.py:4-11)."""
    result = 0.0
    factor = 1.0
    while index > 0:
        factor /= base
        result += factor * (index % base)
        index //= base
    return result


class HaltonSamplePattern:
    """Cycles through ``sample_count`` Halton(2,3) points (camera.py:13-25)."""

    def __init__(self, sample_count: int):
        self.sample_count = int(sample_count)
        self.current_sample = 0

    def next(self) -> tuple[float, float]:
        value = (halton(self.current_sample, 2), halton(self.current_sample, 3))
        self.current_sample += 1
        if self.sample_count != 0:
            self.current_sample %= self.sample_count
        return (value[0] % 1.0, value[1] % 1.0)


def _normalize(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v)


class Camera:
    """Pinhole camera (camera.py:27-117)."""

    CAMERA_RECORD_FLOATS = 16

    def __init__(self):
        self.width = 100
        self.height = 100
        self.position = np.array([1.0, 1.0, 1.0])
        self.target = np.array([0.0, 0.0, 0.0])
        self.up = np.array([0.0, 1.0, 0.0])
        self.fov = 70.0
        self.near_clip_plane = 0.03
        self.far_clip_plane = 1000.0
        self.focal_distance = 10000.0
        self.sample_pattern = HaltonSamplePattern(16)
        self.jitter = (0.0, 0.0)
        self.frame_index = 0
        self.recompute()

    def recompute(self) -> None:
        position = np.asarray(self.position, dtype=np.float64)
        target = np.asarray(self.target, dtype=np.float64)
        up = np.array([0.0, 1.0, 0.0])
        self.aspect_ratio = float(self.width) / float(self.height)
        fwd = _normalize(target - position)
        fov = math.radians(self.fov)
        image_w = fwd * self.focal_distance
        image_u = _normalize(np.cross(image_w, up))
        image_v = _normalize(np.cross(image_u, image_w))
        image_u = image_u * (self.focal_distance * math.tan(fov * 0.5) * self.aspect_ratio)
        image_v = image_v * (self.focal_distance * math.tan(fov * 0.5))
        self.image_u, self.image_v, self.image_w = image_u, image_v, image_w

    def begin_frame(self, width: int, height: int) -> None:
        """Advance the Halton jitter for a new frame (camera.py:87-96)."""
        self.width = int(width)
        self.height = int(height)
        sample = self.sample_pattern.next()
        self.jitter = (sample[0] / self.width, sample[1] / self.height)
        self.recompute()
        self.frame_index += 1

    def pack(self) -> np.ndarray:
        """``float32[16]``: position, jitter_x, image_u, jitter_y, image_v,
        near, image_w, far (the member order of camera.slang:13-20)."""
        rec = np.zeros(self.CAMERA_RECORD_FLOATS, dtype=np.float32)
        rec[0:3] = np.asarray(self.position, dtype=np.float64)
        rec[3] = self.jitter[0]
        rec[4:7] = self.image_u
        rec[7] = self.jitter[1]
        rec[8:11] = self.image_v
        rec[11] = self.near_clip_plane
        rec[12:15] = self.image_w
        rec[15] = self.far_clip_plane
        return rec
