"""The renderer plug-in protocol (renderer.py:6-22 of the reference)."""

from __future__ import annotations

from typing import Protocol

import torch

from .device import CommandEncoder, Device
from .render_data import RenderData
from .scene import Scene


class Renderer(Protocol):
    def initialize(self, device: Device, scene: Scene) -> None:
        ...

    def render(self, command_encoder: CommandEncoder, output: torch.Tensor, frame: int, device: Device,
               scene: Scene, render_data: RenderData) -> None:
        ...

    def setup_ui(self, ui_context, ui_window) -> None:
        ...
