"""Atmosphere lookup tables (atmosphere.py of the reference; Hillaire 2020).

Three generators with the reference's interface: ``generate`` records the
kernel on the command encoder's stream, ``get_texture`` returns the table, a
float32 tensor ``[height, width, 4]`` on the device (the reference's
RGBA32F texture).  The transmittance (256x64) and multiple-scattering (32x32)
tables depend on nothing and are made once; the sky-view table (200x200)
depends on the sun direction.

:class:`SkyAtmosphere` is the part of the reference's ``Scene`` that owns
these generators (scene.py:254-263, 1055-1072): it regenerates the sky-view
table when the sun moves and gives the transmittance at the view point
(``get_sun_transmittance_at_view``, scene.slang:997-1009) that scales sun
next-event estimation.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native


def _sun3(sun_direction) -> np.ndarray:
    d = np.ascontiguousarray(sun_direction, dtype=np.float32).reshape(3)
    if not np.all(np.isfinite(d)) or not np.any(d != 0):
        raise ValueError("sun_direction must be a finite non-zero 3-vector")
    return d


def _f32(a: np.ndarray):
    return a.ctypes.data_as(_native.f32p)


def _check_table(t: torch.Tensor, shape: tuple, name: str, device) -> None:
    if t.dtype != torch.float32 or tuple(t.shape) != shape or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous float32 tensor of shape {shape}")
    if t.device != device.torch_device:
        raise ValueError(f"{name} must live on the generator's device")


class AtmosphereTransmittanceLUT:
    """Transmittance to the top of the atmosphere per (sun zenith cosine,
    height): 256 x 64 (atmosphere.py:4-63)."""

    LUT_WIDTH = 256
    LUT_HEIGHT = 64

    def __init__(self, device):
        self.device = device
        self.transmittance_lut = device.zeros((self.LUT_HEIGHT, self.LUT_WIDTH, 4))

    def generate(self, command_encoder) -> None:
        _native.check(_native.lib().nan_atmosphere_transmittance_lut(self.transmittance_lut.data_ptr(),
                                                                      command_encoder.stream_handle),
                      "nan_atmosphere_transmittance_lut")

    def get_texture(self) -> torch.Tensor:
        return self.transmittance_lut


class AtmosphereMultiScatteringLUT:
    """Multiple-scattering term Psi_ms per (sun zenith cosine, height):
    32 x 32 (atmosphere.py:66-140)."""

    LUT_WIDTH = 32
    LUT_HEIGHT = 32

    def __init__(self, device):
        self.device = device
        self.multiscatt_lut = device.zeros((self.LUT_HEIGHT, self.LUT_WIDTH, 4))

    def generate(self, command_encoder, transmittance_lut: torch.Tensor) -> None:
        _check_table(transmittance_lut, (64, 256, 4), "transmittance_lut", self.device)
        _native.check(_native.lib().nan_atmosphere_multiscattering_lut(
            transmittance_lut.data_ptr(), self.multiscatt_lut.data_ptr(), command_encoder.stream_handle),
            "nan_atmosphere_multiscattering_lut")

    def get_texture(self) -> torch.Tensor:
        return self.multiscatt_lut


class AtmosphereSkyViewLUT:
    """Sky radiance per (azimuth around the sun, non-linear altitude) seen
    from 200 m above the ground: 200 x 200 (atmosphere.py:143-228)."""

    LUT_WIDTH = 200
    LUT_HEIGHT = 200

    def __init__(self, device):
        self.device = device
        self.sky_view_lut = device.zeros((self.LUT_HEIGHT, self.LUT_WIDTH, 4))

    def generate(self, command_encoder, transmittance_lut: torch.Tensor, multiscatt_lut: torch.Tensor,
                 sun_direction: tuple = (0.0, 0.5, -0.866)) -> None:
        _check_table(transmittance_lut, (64, 256, 4), "transmittance_lut", self.device)
        _check_table(multiscatt_lut, (32, 32, 4), "multiscatt_lut", self.device)
        sun = _sun3(sun_direction)
        _native.check(_native.lib().nan_atmosphere_sky_view_lut(
            transmittance_lut.data_ptr(), multiscatt_lut.data_ptr(), _f32(sun), self.sky_view_lut.data_ptr(),
            command_encoder.stream_handle), "nan_atmosphere_sky_view_lut")

    def get_texture(self) -> torch.Tensor:
        return self.sky_view_lut


class SkyAtmosphere:
    """The three tables for one sun direction (scene.py:254-263, 1055-1072)."""

    def __init__(self, device, sun_direction=(0.0, 0.5, 0.866)):
        self.device = device
        self.transmittance_lut_gen = AtmosphereTransmittanceLUT(device)
        self.multiscatt_lut_gen = AtmosphereMultiScatteringLUT(device)
        self.sky_view_lut_gen = AtmosphereSkyViewLUT(device)
        self._view_transmittance = device.zeros((4,))
        enc = device.create_command_encoder()
        self.transmittance_lut_gen.generate(enc)
        self.multiscatt_lut_gen.generate(enc, self.transmittance_lut_gen.get_texture())
        self.sun_direction = None
        self.set_sun_direction(sun_direction)

    def set_sun_direction(self, sun_direction) -> None:
        """Regenerates the sky-view table when the direction changes."""
        sun = tuple(float(v) for v in _sun3(sun_direction))
        if sun == self.sun_direction:
            return
        enc = self.device.create_command_encoder()
        self.sky_view_lut_gen.generate(enc, self.transmittance_lut_gen.get_texture(),
                                       self.multiscatt_lut_gen.get_texture(), sun)
        _native.check(_native.lib().nan_atmosphere_view_transmittance(
            self.transmittance_lut_gen.get_texture().data_ptr(), _f32(np.asarray(sun, dtype=np.float32)),
            self._view_transmittance.data_ptr(), enc.stream_handle), "nan_atmosphere_view_transmittance")
        self.sun_direction = sun

    @property
    def sky_view_lut(self) -> torch.Tensor:
        return self.sky_view_lut_gen.get_texture()

    def get_sun_transmittance_at_view(self) -> tuple:
        """Transmittance from the view point toward the sun (rgb floats)."""
        self.device.wait_for_idle()
        t = self._view_transmittance.cpu().numpy()
        return (float(t[0]), float(t[1]), float(t[2]))

    def sample_sky_lut(self, lighting, dirs: torch.Tensor) -> torch.Tensor:
        """Scene.sample_sky_lut for directions float32[n, 4] on the device
        (``lighting``: a Lighting record) -> float32[n, 4]."""
        if dirs.dtype != torch.float32 or dirs.dim() != 2 or dirs.shape[1] != 4 or not dirs.is_contiguous():
            raise ValueError("dirs must be a contiguous float32[n, 4] tensor")
        out = torch.empty_like(dirs)
        rec = np.ascontiguousarray(lighting.pack(), dtype=np.float32)
        _native.check(_native.lib().nan_sky_view_lookup(self.sky_view_lut.data_ptr(), _f32(rec), dirs.data_ptr(),
                                                        dirs.shape[0], out.data_ptr(), self.device.stream_handle),
                      "nan_sky_view_lookup")
        return out
