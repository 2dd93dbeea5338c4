"""Name-keyed lazy cache of device resources (render_data.py:32-152 of the
reference).  A resource is recreated whenever its descriptor changes.
Buffers are 1-D byte-sized tensors viewed as requested; textures are
float32[H, W, 4] tensors."""

from __future__ import annotations

from dataclasses import dataclass

import torch

from .device import Device


@dataclass
class _Record:
    tensor: torch.Tensor
    key: tuple


class RenderData:
    def __init__(self, device: Device):
        self.device = device
        self._textures: dict[str, _Record] = {}
        self._buffers: dict[str, _Record] = {}

    def get_texture(self, name: str, *, width: int, height: int, channels: int = 4,
                    dtype: torch.dtype = torch.float32, **kwargs) -> torch.Tensor:
        key = (int(width), int(height), int(channels), dtype, tuple(sorted(kwargs.items())))
        rec = self._textures.get(name)
        if rec is not None and rec.key == key:
            return rec.tensor
        t = torch.zeros((int(height), int(width), int(channels)), dtype=dtype, device=self.device.torch_device)
        self._textures[name] = _Record(t, key)
        return t

    def get_buffer(self, name: str, *, size: int | None = None, struct_size: int | None = None,
                   element_count: int | None = None, **kwargs) -> torch.Tensor:
        """Returns a zero-initialised buffer of ``size`` bytes, or
        ``element_count`` records of ``struct_size`` bytes viewed as
        float32[element_count, struct_size/4] when the record is a multiple of
        4 bytes."""
        if size is None and (struct_size is None or element_count is None):
            raise ValueError("get_buffer requires either 'size' or both 'struct_size' and 'element_count'.")
        key = (size, struct_size, element_count, tuple(sorted(kwargs.items())))
        rec = self._buffers.get(name)
        if rec is not None and rec.key == key:
            return rec.tensor
        if size is not None:
            t = torch.zeros((int(size),), dtype=torch.uint8, device=self.device.torch_device)
        elif struct_size % 4 == 0:
            t = torch.zeros((int(element_count), int(struct_size) // 4), dtype=torch.float32,
                            device=self.device.torch_device)
        else:
            t = torch.zeros((int(element_count) * int(struct_size),), dtype=torch.uint8,
                            device=self.device.torch_device)
        self._buffers[name] = _Record(t, key)
        return t

    def release_texture(self, name: str) -> None:
        self._textures.pop(name, None)

    def release_buffer(self, name: str) -> None:
        self._buffers.pop(name, None)

    def has_buffer(self, name: str) -> bool:
        return name in self._buffers

    def clear(self) -> None:
        self._textures.clear()
        self._buffers.clear()
