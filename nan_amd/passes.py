"""Resolve, RGB9E5 freeze, screen accumulation and tone mapping passes.

* :class:`MeshColorsRGB9E5Packer` - mesh_colors_rgb9e5.py:4-29
* :class:`MeshColorsResolve` - mesh_colors_resolve.py:8-73
* :class:`Accumulator` - accumulator.py:9-59
* :class:`ToneMapper` - tone_mapper.py:9-42

Each launches one gfx950 kernel of libnan_hip on the encoder's stream.
Images are float32[H, W, 4] device tensors.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native
from .device import CommandEncoder, Device
from .mesh_colors import MeshColorsLayout
from .render_data import RenderData
from .scene import Scene
from .texture_space_path_tracer import TextureSpacePathTracer, as_payload_tensor


def _check_image(t: torch.Tensor, name: str) -> None:
    if t.dtype != torch.float32 or t.dim() != 3 or t.shape[2] != 4 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError(f"{name} must be a contiguous float32[H, W, 4] device tensor")


class MeshColorsRGB9E5Packer:
    """Freezes float3 slots into one shared-exponent word per slot."""

    def __init__(self, device: Device):
        self.device = device

    def execute(self, command_encoder: CommandEncoder, source: torch.Tensor, destination: torch.Tensor,
                texel_count: int) -> None:
        texel_count = max(0, int(texel_count))
        if texel_count == 0:
            return
        as_payload_tensor(source, texel_count)
        if destination.numel() * destination.element_size() < 4 * texel_count or not destination.is_contiguous():
            raise ValueError("destination must hold texel_count 32-bit words")
        _native.check(_native.lib().nan_rgb9e5_pack(source.data_ptr(), destination.data_ptr(), texel_count,
                                                    command_encoder.stream_handle), "nan_rgb9e5_pack")

    def unpack(self, command_encoder: CommandEncoder, packed: torch.Tensor, texel_count: int) -> torch.Tensor:
        out = torch.empty((texel_count, 4), dtype=torch.float32, device=packed.device)
        _native.check(_native.lib().nan_rgb9e5_unpack(packed.data_ptr(), out.data_ptr(), texel_count,
                                                      command_encoder.stream_handle), "nan_rgb9e5_unpack")
        return out


class MeshColorsResolve:
    """Primary rays into the camera, shading from the Mesh Colors cache."""

    def __init__(self, device: Device, scene: Scene, layout: MeshColorsLayout,
                 texture_space_path_tracer: TextureSpacePathTracer):
        self.device = device
        self.scene = scene
        self.layout = layout
        self.texture_space_path_tracer = texture_space_path_tracer

    def _run(self, command_encoder, cache, output, frozen: bool, camera=None) -> None:
        _check_image(output, "output")
        cam = camera if camera is not None else self.scene.camera
        rec = np.ascontiguousarray(cam.pack(), dtype=np.float32)
        h, w = output.shape[0], output.shape[1]
        _native.check(
            _native.lib().nan_resolve(self.texture_space_path_tracer.handle, cache.data_ptr(), 1 if frozen else 0,
                                      rec.ctypes.data_as(_native.f32p), output.data_ptr(), w, h,
                                      command_encoder.stream_handle), "nan_resolve")

    def execute(self, command_encoder: CommandEncoder, mesh_colors: torch.Tensor, output: torch.Tensor,
                camera=None) -> None:
        as_payload_tensor(mesh_colors, self.layout.total_payload_count)
        self._run(command_encoder, mesh_colors, output, False, camera)

    def execute_frozen(self, command_encoder: CommandEncoder, mesh_colors: torch.Tensor, output: torch.Tensor,
                       camera=None) -> None:
        if mesh_colors.numel() * mesh_colors.element_size() < 4 * self.layout.total_payload_count:
            raise ValueError("frozen cache holds fewer words than payload slots")
        self._run(command_encoder, mesh_colors, output, True, camera)


class Accumulator:
    """Running screen mean, reset on camera moves (accumulator.py:9-59)."""

    def __init__(self, device: Device, resource_key: str = "accumulator.history"):
        self.device = device
        self.resource_key = resource_key

    def execute(self, command_encoder: CommandEncoder, render_data: RenderData, input: torch.Tensor,
                output: torch.Tensor, reset: bool = False) -> None:
        _check_image(input, "input")
        _check_image(output, "output")
        accum = render_data.get_texture(self.resource_key, width=input.shape[1], height=input.shape[0])
        n = input.shape[0] * input.shape[1]
        _native.check(_native.lib().nan_accumulate(input.data_ptr(), accum.data_ptr(), output.data_ptr(), n,
                                                   1 if reset else 0, command_encoder.stream_handle),
                      "nan_accumulate")


class ToneMapper:
    """LPM tone map (hdrMax 12, contrast 1) then Rec.709 (tone_mapper.py:9-42)."""

    def __init__(self, device: Device, exposure: float = 0.0):
        self.device = device
        self.exposure = float(exposure)

    def execute(self, command_encoder: CommandEncoder, input: torch.Tensor, output: torch.Tensor) -> None:
        _check_image(input, "input")
        _check_image(output, "output")
        n = input.shape[0] * input.shape[1]
        _native.check(_native.lib().nan_tonemap(input.data_ptr(), output.data_ptr(), n, float(self.exposure),
                                                command_encoder.stream_handle), "nan_tonemap")
