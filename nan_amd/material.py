"""Materials: the subset of the reference's PBR material the texture-space
tracer reads (material.py:27-67).

The tracer uses ``base_color`` as the Lambert albedo, ``emissive`` as the
emitted radiance and the ``double_sided`` flag, which both selects the
back-slot allocation in the Mesh Colors layout (mesh_colors.py:208-209)
and flips normals toward the viewer during shading (scene.slang:373-382).
Textured materials are outside this path.
"""

from __future__ import annotations

import numpy as np

ALPHA_MODE_OPAQUE = 0
ALPHA_MODE_MASK = 1
ALPHA_MODE_BLEND = 2
MATERIAL_FLAG_DOUBLE_SIDED = 1 << 0

# Packed material record consumed by the native library: 8 x 32-bit words,
# base_color.rgb (f32), flags (u32), emissive.rgb (f32), alpha_mode (u32).
MATERIAL_RECORD_WORDS = 8


class Material:
    """Constant-valued material (material.py:27-67)."""

    def __init__(
        self,
        base_color=(0.5, 0.5, 0.5),
        emissive=(0.0, 0.0, 0.0),
        specular_color=(1.0, 1.0, 1.0),
        roughness: float = 0.5,
        metallic: float = 0.0,
        alpha_mode: int = ALPHA_MODE_OPAQUE,
        alpha_cutoff: float = 0.5,
        double_sided: bool = False,
    ):
        self.base_color = tuple(float(x) for x in _as3(base_color))
        self.emissive = tuple(float(x) for x in _as3(emissive))
        self.specular_color = tuple(float(x) for x in _as3(specular_color))
        self.roughness = float(roughness)
        self.metallic = float(metallic)
        if alpha_mode != ALPHA_MODE_OPAQUE:
            raise ValueError("only opaque materials are supported on the texture-space path")
        self.alpha_mode = int(alpha_mode)
        self.alpha_cutoff = float(alpha_cutoff)
        self.double_sided = bool(double_sided)
        self.flags = MATERIAL_FLAG_DOUBLE_SIDED if self.double_sided else 0
        self.shading_model = 65535

    def pack(self) -> np.ndarray:
        """One 32-byte record as ``uint32[8]``."""
        rec = np.zeros(MATERIAL_RECORD_WORDS, dtype=np.uint32)
        f = rec.view(np.float32)
        f[0:3] = self.base_color
        rec[3] = self.flags
        f[4:7] = self.emissive
        rec[7] = self.alpha_mode
        return rec


def _as3(value):
    if np.isscalar(value):
        return (value, value, value)
    value = tuple(value)
    if len(value) != 3:
        raise ValueError("expected a scalar or 3 components")
    return value
