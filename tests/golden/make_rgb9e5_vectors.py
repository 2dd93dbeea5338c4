"""Float64 RGB9E5 codec restated from the reference's own test
(test_mesh_colors_rgb9e5.py:15-50): the definition the oracle and the
kernels must reproduce bit for bit.

CPU-only: the CPU tests pin the oracle against this definition directly,
and running this file writes the fixture the GPU tests read instead
(``rgb9e5_vectors.npz`` beside it: the reference's seven vectors and the
words this definition assigns them).  The file itself does not travel to
GPU machines.

    python tests/golden/make_rgb9e5_vectors.py
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np

RGB9E5_MAX = 65_408.0


def pack_rgb9e5_reference(value) -> int:
    color = np.asarray(value, dtype=np.float64).copy()
    color = np.nan_to_num(color, nan=0.0, posinf=RGB9E5_MAX, neginf=0.0)
    color = np.clip(color, 0.0, RGB9E5_MAX)
    maximum = float(np.max(color))
    if maximum < 2.0 ** -25:
        return 0
    exponent = max(-15, math.floor(math.log2(maximum)) + 1)
    shared = exponent + 15
    scale = 2.0 ** (shared - 24)
    mantissa = np.floor(color / scale + 0.5).astype(np.uint32)
    if int(np.max(mantissa)) == 512 and shared < 31:
        shared += 1
        scale *= 2.0
        mantissa = np.floor(color / scale + 0.5).astype(np.uint32)
    mantissa = np.minimum(mantissa, 511)
    return int(mantissa[0] | (mantissa[1] << np.uint32(9)) | (mantissa[2] << np.uint32(18))
               | (np.uint32(shared) << np.uint32(27)))


def pack_rgb9e5_reference_array(values: np.ndarray) -> np.ndarray:
    """Vectorised form of :func:`pack_rgb9e5_reference` over float32[N,3]
    (same float64 operations, element-wise)."""
    c = np.asarray(values, dtype=np.float64).reshape(-1, 3).copy()
    c = np.nan_to_num(c, nan=0.0, posinf=RGB9E5_MAX, neginf=0.0)
    c = np.clip(c, 0.0, RGB9E5_MAX)
    m = c.max(axis=1)
    small = m < 2.0 ** -25
    with np.errstate(divide="ignore"):
        e = np.floor(np.log2(np.where(small, 1.0, m))).astype(np.int64) + 1
    e = np.maximum(e, -15)
    shared = e + 15
    scale = np.ldexp(1.0, (shared - 24).astype(np.int64))
    mant = np.floor(c / scale[:, None] + 0.5)
    bump = (mant.max(axis=1) == 512) & (shared < 31)
    shared = np.where(bump, shared + 1, shared)
    scale = np.where(bump, scale * 2.0, scale)
    mant = np.floor(c / scale[:, None] + 0.5)
    mant = np.minimum(mant, 511).astype(np.uint64)
    out = mant[:, 0] | (mant[:, 1] << 9) | (mant[:, 2] << 18) | (shared.astype(np.uint64) << 27)
    return np.where(small, 0, out).astype(np.uint32)


def unpack_rgb9e5_reference(packed: int) -> np.ndarray:
    scale = 2.0 ** ((int(packed) >> 27) - 24)
    return np.asarray([int(packed) & 0x1FF, (int(packed) >> 9) & 0x1FF, (int(packed) >> 18) & 0x1FF],
                      dtype=np.float32) * scale


# The reference's seven vectors (test_mesh_colors_rgb9e5.py:61-72) and the
# words the float64 definition assigns them.
REFERENCE_VECTORS = np.asarray(
    [
        [0.0, 0.0, 0.0],
        [1.0, 1.0, 1.0],
        [0.25, 0.5, 1.0],
        [100.0, 10.0, 1.0],
        [RGB9E5_MAX, RGB9E5_MAX, RGB9E5_MAX],
        [-1.0, np.nan, np.inf],
        [2.0 ** -24, 2.0 ** -20, 2.0 ** -16],
    ],
    dtype=np.float32,
)
REFERENCE_WORDS = np.asarray(
    [0x0, 0x84020100, 0x84010040, 0xB0105190, 0xFFFFFFFF, 0xFFFC0000, 0x4002001], dtype=np.uint32)


FIXTURE = Path(__file__).resolve().parent / "rgb9e5_vectors.npz"


def main() -> None:
    words = pack_rgb9e5_reference_array(REFERENCE_VECTORS)
    # the words transcribed above agree with the definition
    assert (words == REFERENCE_WORDS).all()
    np.savez(FIXTURE, vectors=REFERENCE_VECTORS, words=words)
    print(f"wrote {FIXTURE}")


if __name__ == "__main__":
    main()
