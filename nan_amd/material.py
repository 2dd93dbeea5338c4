"""Materials: the subset of the reference's PBR material the texture-space
tracer reads (material.py:27-67; MaterialDesc in scene.slang).

The tracer shades Lambert surfaces: ``base_color`` (or the base-colour
texture, which replaces it, scene.slang:313-317) is the albedo,
``emissive`` (times the emissive texture, scene.slang:319-322) the emitted
radiance, and the normal texture tilts the shading normal
(scene.slang:342-368).  ``double_sided`` selects the back-slot allocation
of the Mesh Colors layout (mesh_colors.py:208-209) and flips normals
toward the viewer (scene.slang:373-382).  Non-opaque alpha modes are alpha
tested on every candidate hit (scene.slang:124-149): mask keeps a hit whose
opacity (the base-colour texture's alpha, 1 without one) reaches
``alpha_cutoff``, blend keeps it with probability equal to the opacity.
Roughness, metallic and specular inputs are kept for the caller but the
Lambert path reads none of them.

A texture is a file path, encoded image bytes or an array (see
``textures.load_image``).
"""

from __future__ import annotations

import numpy as np

from .textures import NO_TEXTURE

ALPHA_MODE_OPAQUE = 0
ALPHA_MODE_MASK = 1
ALPHA_MODE_BLEND = 2
MATERIAL_FLAG_DOUBLE_SIDED = 1 << 0

# Packed material record consumed by the native library: 12 x 32-bit words,
# base_color.rgb (f32), flags (u32), emissive.rgb (f32), alpha_mode (u32),
# base-colour, emissive and normal texture index (u32, NO_TEXTURE: none),
# alpha_cutoff (f32).
MATERIAL_RECORD_WORDS = 12

# texture slots the path samples, and whether each holds sRGB colour
# (texture_manager.py:93-94)
TEXTURE_SLOTS = (("base_color_texture", True), ("emissive_texture", True), ("normal_texture", False))


class Material:
    """PBR material (material.py:27-67) as the texture-space tracer uses it."""

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
        base_color_texture=None,
        emissive_texture=None,
        normal_texture=None,
        roughness_texture=None,
        metallic_texture=None,
        specular_color_texture=None,
    ):
        self.base_color = tuple(float(x) for x in _as3(base_color))
        self.emissive = tuple(float(x) for x in _as3(emissive))
        self.specular_color = tuple(float(x) for x in _as3(specular_color))
        self.roughness = float(roughness)
        self.metallic = float(metallic)
        if alpha_mode not in (ALPHA_MODE_OPAQUE, ALPHA_MODE_MASK, ALPHA_MODE_BLEND):
            raise ValueError(f"unknown alpha_mode {alpha_mode}")
        self.alpha_mode = int(alpha_mode)
        self.alpha_cutoff = float(alpha_cutoff)
        self.double_sided = bool(double_sided)
        self.flags = MATERIAL_FLAG_DOUBLE_SIDED if self.double_sided else 0
        self.shading_model = 65535
        self.base_color_texture = base_color_texture
        self.emissive_texture = emissive_texture
        self.normal_texture = normal_texture
        self.roughness_texture = roughness_texture
        self.metallic_texture = metallic_texture
        self.specular_color_texture = specular_color_texture

    @property
    def textured(self) -> bool:
        return any(getattr(self, name) is not None for name, _ in TEXTURE_SLOTS)

    def pack(self, texture_ids=(NO_TEXTURE, NO_TEXTURE, NO_TEXTURE)) -> np.ndarray:
        """One 48-byte record as ``uint32[12]``; ``texture_ids`` are the
        scene's indices of the base-colour, emissive and normal textures."""
        rec = np.zeros(MATERIAL_RECORD_WORDS, dtype=np.uint32)
        f = rec.view(np.float32)
        f[0:3] = self.base_color
        rec[3] = self.flags
        f[4:7] = self.emissive
        rec[7] = self.alpha_mode
        rec[8:11] = texture_ids
        f[11] = self.alpha_cutoff
        return rec


def _as3(value):
    if np.isscalar(value):
        return (value, value, value)
    value = tuple(value)
    if len(value) != 3:
        raise ValueError("expected a scalar or 3 components")
    return value
