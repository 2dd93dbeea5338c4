"""``--renderer texture-space``: the renderer that owns the Mesh Colors
cache and chains trace -> resolve -> accumulate -> tone map
(texture_space_path_tracing_renderer.py:21-417 of the reference).

The state machine is the reference's: reset / pause / save-to-RGB9E5 /
frozen resolve, camera-move resets, and the save failure path that resets
the texture space.  UI controls are optional attributes that stay ``None``
in headless use (their defaults then apply, as in the reference).  Before
a save, the Mesh Colors surface filter runs its ping-pong passes
(texture_space_path_tracing_renderer.py:181-231); with no ``surface_filter``
attached the save packs the working cache directly.
"""

from __future__ import annotations

import math

import numpy as np

import torch

from .device import CommandEncoder, Device
from .mesh_colors import MESH_COLORS_PAYLOAD_SIZE, MeshColorsLayout
from .mesh_colors_surface_filter import MeshColorsSurfaceFilter
from .passes import Accumulator, MeshColorsResolve, MeshColorsRGB9E5Packer, ToneMapper
from .render_data import RenderData
from .scene import Scene
from .texture_space_path_tracer import TextureSpacePathTracer

MESH_COLORS_BUFFER_KEY = "texture_space_renderer.mesh_colors"
MESH_COLORS_FILTER_SCRATCH_KEY = "texture_space_renderer.mesh_colors_filter_scratch"
MESH_COLORS_RGB9E5_KEY = "texture_space_renderer.mesh_colors_rgb9e5"


class TextureSpacePathTracingRenderer:
    def __init__(self, *, texels_per_unit: float = 16.0, min_resolution: int = 4, max_resolution: int = 64,
                 max_texels: int = 16_777_216, samples_per_texel: int = 1, max_bounces: int = 3):
        self.texels_per_unit = texels_per_unit
        self.min_resolution = min_resolution
        self.max_resolution = max_resolution
        self.max_texels = max_texels
        self.samples_per_texel = samples_per_texel
        self.max_bounces = max_bounces
        self.reset_accumulator = True
        self.reset_texture_space = True
        self.texture_iteration = 0
        self.use_accum_check_box = None
        self.pause_check_box = None
        self.filter_before_save_check_box = None
        self.filter_pass_count_slider = None
        self.filter_spatial_sigma_slider = None
        self.filter_normal_angle_slider = None
        self.save_rgb9e5_button = None
        self.status_text = None
        self._output_size = None
        self._save_requested = False
        self._frozen_rgb9e5 = None
        self._release_frozen_buffer = False
        self._save_error = None
        self.surface_filter = None

    # -- setup ---------------------------------------------------------------
    def initialize(self, device: Device, scene: Scene) -> None:
        self.device = device
        self.scene = scene
        self.layout = MeshColorsLayout.build(
            scene.scene_node,
            texels_per_unit=self.texels_per_unit,
            min_resolution=self.min_resolution,
            max_resolution=self.max_resolution,
            max_total_texels=self.max_texels,
        )
        self.texture_space_path_tracer = TextureSpacePathTracer(
            device, scene, self.layout, samples_per_texel=self.samples_per_texel, max_bounces=self.max_bounces)
        self.resolve = MeshColorsResolve(device, scene, self.layout, self.texture_space_path_tracer)
        self.surface_filter = MeshColorsSurfaceFilter(
            device, scene, self.layout,
            face_infos_buffer=self.texture_space_path_tracer.face_infos_buffer,
            instance_infos_buffer=self.texture_space_path_tracer.instance_infos_buffer,
            side_infos_buffer=self.texture_space_path_tracer.side_infos_buffer,
            adjacency_infos_buffer=self.texture_space_path_tracer.adjacency_infos_buffer,
            texture_space_path_tracer=self.texture_space_path_tracer)
        self.rgb9e5_packer = MeshColorsRGB9E5Packer(device)
        self.accumulator = Accumulator(device, resource_key="texture_space_renderer.accumulator_history")
        self.tone_mapper = ToneMapper(device)
        print(self.layout.summary())

    def on_camera_move(self, data) -> None:
        self.reset_accumulator = True
        if isinstance(data, dict) and data.get("texture_space"):
            self._discard_rgb9e5()
            self.reset_texture_space = True
            self.texture_iteration = 0

    # -- control values (UI or defaults) -------------------------------------
    def _use_accumulation(self) -> bool:
        return True if self.use_accum_check_box is None else bool(self.use_accum_check_box.value)

    def _is_paused(self) -> bool:
        return False if self.pause_check_box is None else bool(self.pause_check_box.value)

    def _filter_before_save(self) -> bool:
        return True if self.filter_before_save_check_box is None else bool(self.filter_before_save_check_box.value)

    def _filter_pass_count(self) -> int:
        return 5 if self.filter_pass_count_slider is None else max(0, int(self.filter_pass_count_slider.value))

    def _filter_spatial_sigma(self) -> float:
        return 1.0 if self.filter_spatial_sigma_slider is None else max(0.05, float(self.filter_spatial_sigma_slider.value))

    def _filter_normal_sigma_radians(self) -> float:
        degrees = 30.0 if self.filter_normal_angle_slider is None else max(0.0, float(self.filter_normal_angle_slider.value))
        return math.radians(degrees)

    def _filtering_enabled(self) -> bool:
        return self._filter_before_save() and self._filter_pass_count() > 0 and self.surface_filter is not None

    # -- state transitions ---------------------------------------------------
    def _request_texture_reset(self) -> None:
        self._discard_rgb9e5()
        self.reset_texture_space = True
        self.reset_accumulator = True
        self.texture_iteration = 0
        if self.pause_check_box is not None:
            self.pause_check_box.value = False

    def _discard_rgb9e5(self) -> None:
        self._save_requested = False
        self._frozen_rgb9e5 = None
        self._release_frozen_buffer = True
        self._save_error = None
        if self.save_rgb9e5_button is not None:
            self.save_rgb9e5_button.enabled = True

    # -- on-disk frozen cache (no reference counterpart: it freezes in memory) --
    def write_frozen_cache(self, path) -> None:
        """Writes the current RGB9E5 frozen cache with its layout tables
        (:mod:`nan_amd.frozen_cache`)."""
        from .frozen_cache import save_frozen_cache

        if self._frozen_rgb9e5 is None:
            raise RuntimeError("no frozen RGB9E5 cache to write; request a save first")
        save_frozen_cache(path, self.layout, self._frozen_rgb9e5)

    def read_frozen_cache(self, path, render_data: RenderData) -> None:
        """Loads a frozen cache written for this renderer's layout and shows
        it (frozen resolve) until the texture space is reset."""
        from .frozen_cache import load_frozen_cache

        words, _ = load_frozen_cache(path, self.layout)
        packed = render_data.get_buffer(MESH_COLORS_RGB9E5_KEY, struct_size=4,
                                        element_count=self.layout.total_payload_count)
        packed.view(-1).copy_(torch.from_numpy(words.view(np.float32)))
        self._frozen_rgb9e5 = packed
        self._save_requested = False
        self._save_error = None
        self.reset_accumulator = True

    def _request_rgb9e5_save(self) -> None:
        if self._save_requested or self._frozen_rgb9e5 is not None:
            return
        if self.texture_iteration < 1:
            self._save_error = "RGB9E5 save requires at least one texture iteration"
            return
        self._save_requested = True
        self._save_error = None
        if self.pause_check_box is not None:
            self.pause_check_box.value = True
        if self.save_rgb9e5_button is not None:
            self.save_rgb9e5_button.enabled = False

    def _execute_rgb9e5_save(self, command_encoder, mesh_colors, filter_scratch, packed) -> bool:
        """Optional ping-pong filter passes, then the RGB9E5 pack
        (texture_space_path_tracing_renderer.py:181-231)."""
        self._save_requested = False
        try:
            source = mesh_colors
            pass_count = self._filter_pass_count()
            if self._filter_before_save() and pass_count > 0 and getattr(self, "surface_filter", None) is not None:
                if filter_scratch is None:
                    raise RuntimeError("Mesh Colors filter scratch buffer is unavailable")
                destination = filter_scratch
                for _ in range(pass_count):
                    self.surface_filter.execute(
                        command_encoder, source, destination,
                        spatial_sigma=self._filter_spatial_sigma(),
                        normal_sigma_radians=self._filter_normal_sigma_radians(),
                    )
                    source, destination = destination, source
            self.rgb9e5_packer.execute(command_encoder, source, packed, self.layout.total_payload_count)
            self._frozen_rgb9e5 = packed
            self._save_error = None
            self.reset_accumulator = True
            return True
        except Exception as exc:  # the reference resets the cache on any failure
            self._frozen_rgb9e5 = None
            self._release_frozen_buffer = True
            self._save_error = str(exc)
            self.reset_texture_space = True
            self.texture_iteration = 0
            if self.save_rgb9e5_button is not None:
                self.save_rgb9e5_button.enabled = True
            return False

    def cache_state(self) -> dict:
        """Where the cache stands, for a headless caller: iterations done,
        slot counts, whether it is frozen (and its RGB9E5 size) or still
        iterating (and its working size), and the last save error."""
        slots = self.layout.total_payload_count
        frozen = self._frozen_rgb9e5 is not None
        return {
            "iterations": self.texture_iteration,
            "surface_slots": self.layout.total_surface_texels,
            "back_slots": slots - self.layout.total_surface_texels,
            "frozen": frozen,
            "save_pending": bool(self._save_requested),
            "save_error": self._save_error,
            "cache_bytes": slots * (4 if frozen else MESH_COLORS_PAYLOAD_SIZE),
        }

    # -- per frame -----------------------------------------------------------
    def render(self, command_encoder: CommandEncoder, output: torch.Tensor, frame: int, device: Device,
               scene: Scene, render_data: RenderData) -> None:
        height, width = int(output.shape[0]), int(output.shape[1])
        if (width, height) != self._output_size:
            self._output_size = (width, height)
            self.reset_accumulator = True
        if self._release_frozen_buffer:
            render_data.release_buffer(MESH_COLORS_RGB9E5_KEY)
            self._release_frozen_buffer = False

        mesh_colors = None
        if self._frozen_rgb9e5 is None:
            mesh_colors = render_data.get_buffer(MESH_COLORS_BUFFER_KEY, struct_size=MESH_COLORS_PAYLOAD_SIZE,
                                                 element_count=self.layout.total_payload_count)
        if self._save_requested and mesh_colors is not None:
            filter_scratch = None
            if self._filtering_enabled():
                filter_scratch = render_data.get_buffer(MESH_COLORS_FILTER_SCRATCH_KEY,
                                                        struct_size=MESH_COLORS_PAYLOAD_SIZE,
                                                        element_count=self.layout.total_payload_count)
            packed = render_data.get_buffer(MESH_COLORS_RGB9E5_KEY, struct_size=4,
                                            element_count=self.layout.total_payload_count)
            if self._execute_rgb9e5_save(command_encoder, mesh_colors, filter_scratch, packed):
                render_data.release_buffer(MESH_COLORS_BUFFER_KEY)
                render_data.release_buffer(MESH_COLORS_FILTER_SCRATCH_KEY)

        if (self._frozen_rgb9e5 is None and not self._save_requested
                and (not self._is_paused() or self.reset_texture_space)):
            resetting = self.reset_texture_space
            self.texture_space_path_tracer.execute(command_encoder, mesh_colors, self.texture_iteration,
                                                   reset=resetting)
            self.texture_iteration += 1
            self.reset_texture_space = False
            if resetting:
                self.reset_accumulator = True

        resolve_texture = render_data.get_texture("texture_space_renderer.resolve_texture", width=width,
                                                  height=height)
        accum_texture = render_data.get_texture("texture_space_renderer.accum_texture", width=width, height=height)
        if self._frozen_rgb9e5 is not None:
            self.resolve.execute_frozen(command_encoder, self._frozen_rgb9e5, resolve_texture)
        else:
            self.resolve.execute(command_encoder, mesh_colors, resolve_texture)
        self.accumulator.execute(command_encoder, render_data, resolve_texture, accum_texture, self.reset_accumulator)
        self.tone_mapper.execute(command_encoder, accum_texture if self._use_accumulation() else resolve_texture,
                                 output)
        self.reset_accumulator = False

    def setup_ui(self, ui_context=None, ui_window=None) -> None:
        """Headless build: no UI toolkit; controls keep their defaults."""
