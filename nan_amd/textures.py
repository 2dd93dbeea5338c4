"""Texture images for the texture-space tracer's materials.

The reference samples base-colour, emissive and normal textures through
bindless handles (scene.slang:313-368) loaded by its TextureManager
(texture_manager.py:70-160; base colour and emissive as sRGB).  Here a
material names each texture as a file path, encoded image bytes (a PNG or
JPEG inside a GLB), or an array; ``load_image`` turns any of them into
RGBA8 rows (row 0 first as stored, R in the lowest byte once packed), and
``pack_textures`` lays every texture of a scene out for the native library:

* ``texels``   uint32[N]     RGBA8 words of all textures, row-major each
* ``textures`` uint32[T, 4]  first texel, width, height, flags (bit 0: sRGB colour)
"""

from __future__ import annotations

import io
from pathlib import Path

import numpy as np

TEXTURE_FLAG_SRGB = 1 << 0
NO_TEXTURE = 0xFFFFFFFF


def load_image(src) -> np.ndarray:
    """RGBA8 pixels ``uint8[H, W, 4]`` of a texture source: an array
    (uint8, or float in [0, 1]; 1 to 4 channels, grey expands to RGB, a
    missing alpha is 255), encoded image bytes, or a file path."""
    if isinstance(src, np.ndarray):
        a = src
        if a.dtype != np.uint8:
            a = np.clip(np.rint(np.asarray(a, dtype=np.float64) * 255.0), 0, 255).astype(np.uint8)
        if a.ndim == 2:
            a = a[:, :, None]
        if a.ndim != 3 or a.shape[2] not in (1, 2, 3, 4) or a.shape[0] < 1 or a.shape[1] < 1:
            raise ValueError(f"texture array of shape {src.shape} is not an image")
        h, w, c = a.shape
        out = np.empty((h, w, 4), dtype=np.uint8)
        if c <= 2:
            out[:, :, 0:3] = a[:, :, 0:1]
            out[:, :, 3] = a[:, :, 1] if c == 2 else 255
        else:
            out[:, :, 0:3] = a[:, :, 0:3]
            out[:, :, 3] = a[:, :, 3] if c == 4 else 255
        return out
    from PIL import Image

    if isinstance(src, (bytes, bytearray, memoryview)):
        img = Image.open(io.BytesIO(bytes(src)))
    else:
        img = Image.open(Path(src))
    return np.ascontiguousarray(np.asarray(img.convert("RGBA"), dtype=np.uint8))


def pack_textures(images: list[np.ndarray], srgb: list[bool]) -> tuple[np.ndarray, np.ndarray]:
    """``texels`` and ``textures`` (see the module docstring) for RGBA8
    images; ``srgb[k]`` marks colour textures whose RGB is sRGB-encoded."""
    table = np.zeros((len(images), 4), dtype=np.uint32)
    words = []
    first = 0
    for k, (img, s) in enumerate(zip(images, srgb)):
        h, w = img.shape[:2]
        px = np.ascontiguousarray(img, dtype=np.uint8).reshape(-1, 4).astype(np.uint32)
        words.append(px[:, 0] | (px[:, 1] << 8) | (px[:, 2] << 16) | (px[:, 3] << 24))
        table[k] = (first, w, h, TEXTURE_FLAG_SRGB if s else 0)
        first += w * h
    texels = np.concatenate(words).astype(np.uint32) if words else np.zeros(0, dtype=np.uint32)
    return texels, table
