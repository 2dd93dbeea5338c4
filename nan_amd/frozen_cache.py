"""On-disk frozen Mesh Colors cache: the RGB9E5 words of a frozen cache
(mesh_colors_rgb9e5_pack.slang:7-16) together with the layout tables that
give them meaning, so a bake can be shipped, checkpointed and resumed.  The
reference freezes only in memory (texture_space_path_tracing_renderer.py:
181-231); this format is new.

File layout (little endian):

    offset  size  field
    0       8     magic  b"NANMCE5\\0"
    8       4     format version (1)
    12      4     flags (bit 0: back slots present)
    16      8     face count F
    24      8     instance count I
    32      8     total surface texels
    40      8     payload slot count N
    48      32    SHA-256 of the face, instance and side tables (as stored)
    80      4     texels_per_unit (float32)
    84      4     min_resolution
    88      4     max_resolution
    92      4     reserved (0)
    96      F*8   face table      uint32[F][2]  (address, resolution)
    ...     I*16  instance table  uint32[I][4]  (face_offset, face_count, texel_offset, texel_count)
    ...     I*8   side table      uint32[I][2]  (back_texel_offset, flags)
    ...     N*4   RGB9E5 words    uint32[N]
    end-32  32    SHA-256 of everything before it

Loading checks the magic, version, sizes, both digests and, when a layout is
given, that its tables hash to the stored value.
"""

from __future__ import annotations

import hashlib
import struct
from pathlib import Path

import numpy as np

from .mesh_colors import MeshColorsLayout

MAGIC = b"NANMCE5\0"
VERSION = 1
HEADER = struct.Struct("<8sII QQQQ 32s fIII")
assert HEADER.size == 96


class FrozenCacheError(ValueError):
    """A frozen-cache file is malformed or does not match the layout."""


def layout_digest(layout: MeshColorsLayout) -> bytes:
    """SHA-256 over the face, instance and side tables."""
    h = hashlib.sha256()
    for table in (layout.face_table, layout.instance_table, layout.side_table):
        h.update(np.ascontiguousarray(table, dtype="<u4").tobytes())
    return h.digest()


def _words(packed) -> np.ndarray:
    if hasattr(packed, "detach"):  # torch tensor (any device)
        packed = packed.detach().cpu().contiguous().view(-1).numpy()
    arr = np.ascontiguousarray(packed).reshape(-1)
    return arr.view(np.uint32) if arr.dtype.itemsize == 4 else arr.astype(np.uint32)


def save_frozen_cache(path, layout: MeshColorsLayout, packed) -> None:
    """Writes the frozen cache (``packed``: RGB9E5 words, numpy or a torch
    tensor of any device, at least ``total_payload_count`` of them)."""
    words = _words(packed)
    n = layout.total_payload_count
    if words.size < n:
        raise ValueError(f"frozen cache holds {words.size} words, the layout has {n} slots")
    words = words[:n].astype("<u4")
    header = HEADER.pack(MAGIC, VERSION, 1 if n > layout.total_surface_texels else 0, len(layout.face_table),
                         len(layout.instance_table), layout.total_surface_texels, n, layout_digest(layout),
                         float(layout.texels_per_unit), int(layout.min_resolution), int(layout.max_resolution), 0)
    body = b"".join([header, np.ascontiguousarray(layout.face_table, dtype="<u4").tobytes(),
                     np.ascontiguousarray(layout.instance_table, dtype="<u4").tobytes(),
                     np.ascontiguousarray(layout.side_table, dtype="<u4").tobytes(), words.tobytes()])
    tmp = Path(str(path) + ".tmp")
    with open(tmp, "wb") as f:
        f.write(body)
        f.write(hashlib.sha256(body).digest())
    tmp.replace(path)


def load_frozen_cache(path, layout: MeshColorsLayout | None = None) -> tuple[np.ndarray, MeshColorsLayout]:
    """Reads a frozen cache; returns (RGB9E5 words uint32[N], the stored
    layout).  With ``layout``, raises FrozenCacheError unless its tables are
    the stored ones."""
    data = Path(path).read_bytes()
    if len(data) < HEADER.size + 32:
        raise FrozenCacheError("file too short for a frozen Mesh Colors cache")
    magic, version, _flags, faces, instances, surface, slots, digest, tpu, min_res, max_res, _ = \
        HEADER.unpack_from(data, 0)
    if magic != MAGIC:
        raise FrozenCacheError("not a frozen Mesh Colors cache (bad magic)")
    if version != VERSION:
        raise FrozenCacheError(f"unsupported frozen-cache version {version}")
    expected = HEADER.size + faces * 8 + instances * 16 + instances * 8 + slots * 4 + 32
    if len(data) != expected:
        raise FrozenCacheError(f"frozen-cache size {len(data)} does not match its header ({expected})")
    if hashlib.sha256(data[:-32]).digest() != data[-32:]:
        raise FrozenCacheError("frozen-cache checksum mismatch")
    off = HEADER.size
    face = np.frombuffer(data, dtype="<u4", count=faces * 2, offset=off).reshape(-1, 2)
    off += faces * 8
    inst = np.frombuffer(data, dtype="<u4", count=instances * 4, offset=off).reshape(-1, 4)
    off += instances * 16
    side = np.frombuffer(data, dtype="<u4", count=instances * 2, offset=off).reshape(-1, 2)
    off += instances * 8
    words = np.frombuffer(data, dtype="<u4", count=slots, offset=off).astype(np.uint32)
    stored = MeshColorsLayout(face.astype(np.uint32), inst.astype(np.uint32), side.astype(np.uint32), surface, slots,
                              tpu, min_res, max_res)
    if layout_digest(stored) != digest:
        raise FrozenCacheError("frozen-cache layout digest mismatch")
    if layout is not None and layout_digest(layout) != digest:
        raise FrozenCacheError("frozen cache was saved for a different Mesh Colors layout")
    return words, stored
