"""GPU scene: uploads a :class:`SceneNode` and builds its BVH
(``Scene.__init__``, scene.py:110-443 of the reference).

The reference builds one BLAS per mesh plus a TLAS through the vendor API
and traverses them with hardware ray queries.  Here the native library
flattens every instance into world-space triangles, builds one binned-SAH
4-wide quantized BVH on the host and uploads it with the shading buffers
(vertices, indices, mesh/instance descriptors, transforms, materials).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .device import Device
from .scene_arrays import flatten_scene
from .scene_node import SceneNode


def _f32(a):
    return a.ctypes.data_as(_native.f32p)


def _u32(a):
    return a.ctypes.data_as(_native.u32p)


class Scene:
    """Scene buffers and BVH on one device."""

    def __init__(self, device: Device, scene_node: SceneNode, *, sah_bins: int = 16, max_leaf_size: int = 4):
        self.device = device
        self.scene_node = scene_node
        self.camera = scene_node.camera
        self.arrays = flatten_scene(scene_node)
        self.atmosphere = None
        self._apply_atmosphere(scene_node.lighting)
        a = self.arrays
        desc = _native.NanSceneDesc(
            _f32(a["vertices"]), a["vertices"].shape[0],
            _u32(a["indices"]), a["indices"].size,
            _u32(a["mesh_descs"]), a["mesh_descs"].shape[0],
            _u32(a["instance_descs"]), a["instance_descs"].shape[0],
            _f32(a["transforms"]), _f32(a["inv_transposes"]), a["transforms"].shape[0],
            _u32(a["materials"]), a["materials"].shape[0],
            _f32(a["lighting"]),
            _u32(a["textures"]), a["textures"].shape[0],
            _u32(a["texels"]), a["texels"].size,
        )
        opts = _native.NanBvhOptions(sah_bins, max_leaf_size, 0)
        handle = ctypes.c_void_p()
        err = _native.err_buffer()
        lib = _native.lib()
        rc = lib.nan_scene_create(ctypes.byref(desc), ctypes.byref(opts), ctypes.byref(handle), err, len(err))
        if rc != 0:
            raise _native.NanError(err.value.decode())
        self.handle = handle
        self.info = _native.NanSceneInfo()
        _native.check(lib.nan_scene_info(self.handle, ctypes.byref(self.info)), "nan_scene_info")
        self.frame_index = 0
        self._attach_sky()
        if a["russian_roulette"]:
            self.set_russian_roulette(True)

    def _apply_atmosphere(self, lighting) -> None:
        """With ``lighting.atmosphere``, (re)generates the tables for its sun
        and writes the view-point transmittance into the lighting record."""
        if not lighting.atmosphere:
            return
        from .atmosphere import SkyAtmosphere

        sun = self.arrays["lighting"][0:3].copy()
        if self.atmosphere is None:
            self.atmosphere = SkyAtmosphere(self.device, sun)
        else:
            self.atmosphere.set_sun_direction(sun)
        self.arrays["lighting"][8:11] = self.atmosphere.get_sun_transmittance_at_view()

    def _attach_sky(self) -> None:
        lut = self.atmosphere.sky_view_lut.data_ptr() \
            if self.atmosphere is not None and self.scene_node.lighting.atmosphere else None
        _native.check(_native.lib().nan_scene_set_sky_view_lut(self.handle, lut), "nan_scene_set_sky_view_lut")

    def close(self) -> None:
        if getattr(self, "handle", None):
            _native.lib().nan_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bounds(self) -> tuple:
        """World-space bounding box (min, max float64[3]) of every instance."""
        from .scene_arrays import world_bounds

        return world_bounds(self.arrays)

    @property
    def triangle_count(self) -> int:
        return int(self.info.triangle_count)

    def update(self) -> None:
        """Per-frame scene update (scene.py:1213-1232): advances the frame
        index that the stochastic alpha test of blend materials hashes."""
        self.set_frame_index(self.frame_index + 1)

    def set_frame_index(self, frame_index: int) -> None:
        """Sets the blend alpha test's frame index (scene.slang:146-148),
        written on the device's stream."""
        self.frame_index = int(frame_index) & 0xFFFFFFFF
        _native.check(_native.lib().nan_scene_set_frame_index(self.handle, self.frame_index,
                                                               self.device.stream_handle),
                      "nan_scene_set_frame_index")

    def set_lighting(self, lighting) -> None:
        self.scene_node.lighting = lighting
        self.arrays["lighting"] = lighting.pack()
        self._apply_atmosphere(lighting)
        _native.check(_native.lib().nan_scene_set_lighting(self.handle, _f32(self.arrays["lighting"])),
                      "nan_scene_set_lighting")
        self._attach_sky()

    def set_russian_roulette(self, enable: bool) -> None:
        """Russian roulette on this scene's texture-space paths: a path that
        would take another bounce survives with probability min(1, max
        throughput component) and is reweighted by its inverse.  An extension:
        the reference's paths always run to max_bounces
        (texture_space_path_tracer.slang:57-84)."""
        self.scene_node.russian_roulette = bool(enable)
        self.arrays["russian_roulette"] = bool(enable)
        _native.check(_native.lib().nan_scene_set_russian_roulette(self.handle, 1 if enable else 0),
                      "nan_scene_set_russian_roulette")

    def trace_rays(self, rays: torch.Tensor, *, occlusion: bool = True):
        """Closest hit (and occlusion) for ``rays`` float32[n, 8] on the
        device.  Returns ``(hits float32[n,4] = t, b1, b2, 0, ids int32[n,2]
        = instance, primitive, occluded int32[n])``; misses have t < 0 and
        ids -1."""
        if rays.dtype != torch.float32 or rays.dim() != 2 or rays.shape[1] != 8 or not rays.is_contiguous():
            raise ValueError("rays must be a contiguous float32[n, 8] tensor")
        if rays.device != self.device.torch_device:
            raise ValueError("rays must live on the scene's device")
        n = rays.shape[0]
        hits = torch.empty((n, 4), dtype=torch.float32, device=rays.device)
        ids = torch.empty((n, 2), dtype=torch.int32, device=rays.device)
        occ = torch.empty((n,), dtype=torch.int32, device=rays.device) if occlusion else None
        _native.check(
            _native.lib().nan_trace_rays(self.handle, rays.data_ptr(), n, hits.data_ptr(), ids.data_ptr(),
                                         occ.data_ptr() if occ is not None else None, self.device.stream_handle),
            "nan_trace_rays")
        return hits, ids, occ
