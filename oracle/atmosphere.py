"""CPU oracle of the atmosphere tables and the sky-view lookup (test
infrastructure only; wraps oracle/atmosphere_oracle.c and the lookup in
oracle/nan_oracle.c).  Tables are float32 arrays [height, width, 4]."""

from __future__ import annotations

import numpy as np

from . import _p, f32p, liboracle

TRANSMITTANCE_SHAPE = (64, 256, 4)
MULTISCATTERING_SHAPE = (32, 32, 4)
SKY_VIEW_SHAPE = (200, 200, 4)


def _f3(v) -> np.ndarray:
    return np.ascontiguousarray(v, dtype=np.float32).reshape(3)


def scattering(pos) -> tuple[np.ndarray, float, np.ndarray]:
    """getScatteringValues: (rayleigh.rgb, mie, extinction.rgb) per megameter."""
    out = np.zeros(7, dtype=np.float32)
    liboracle().orc_atmo_scattering(_p(_f3(pos), f32p), _p(out, f32p))
    return out[0:3], float(out[3]), out[4:7]


def sphere_hit(origin, direction, radius: float) -> float:
    return float(liboracle().orc_atmo_sphere_hit(_p(_f3(origin), f32p), _p(_f3(direction), f32p), radius))


def sun_transmittance(pos, direction) -> np.ndarray:
    out = np.zeros(3, dtype=np.float32)
    liboracle().orc_atmo_sun_transmittance(_p(_f3(pos), f32p), _p(_f3(direction), f32p), _p(out, f32p))
    return out


def transmittance_lut() -> np.ndarray:
    out = np.zeros(TRANSMITTANCE_SHAPE, dtype=np.float32)
    liboracle().orc_atmo_transmittance_lut(_p(out, f32p))
    return out


def multiscattering_lut(transmittance: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(transmittance, dtype=np.float32)
    out = np.zeros(MULTISCATTERING_SHAPE, dtype=np.float32)
    liboracle().orc_atmo_multiscattering_lut(_p(t, f32p), _p(out, f32p))
    return out


def multiscattering_texel(transmittance: np.ndarray, px: int, py: int, directions: int, steps: int) -> np.ndarray:
    """One texel of the multiple-scattering table computed with a
    directions x directions rule and `steps` ray-march steps (the table uses
    the reference's 8 and 20); float32[3]."""
    t = np.ascontiguousarray(transmittance, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    liboracle().orc_atmo_multiscattering_texel(_p(t, f32p), int(px), int(py), int(directions), int(steps),
                                               _p(out, f32p))
    return out


def sky_view_lut(transmittance: np.ndarray, multiscattering: np.ndarray, sun_direction) -> np.ndarray:
    t = np.ascontiguousarray(transmittance, dtype=np.float32)
    m = np.ascontiguousarray(multiscattering, dtype=np.float32)
    out = np.zeros(SKY_VIEW_SHAPE, dtype=np.float32)
    liboracle().orc_atmo_sky_view_lut(_p(t, f32p), _p(m, f32p), _p(_f3(sun_direction), f32p), _p(out, f32p))
    return out


def view_transmittance(transmittance: np.ndarray, sun_direction) -> np.ndarray:
    t = np.ascontiguousarray(transmittance, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    liboracle().orc_atmo_view_transmittance(_p(t, f32p), _p(_f3(sun_direction), f32p), _p(out, f32p))
    return out


def sky_view_lookup(sky_view: np.ndarray, lighting16: np.ndarray, dirs: np.ndarray) -> np.ndarray:
    """Scene.sample_sky_lut for directions float32[n, 3 or 4] -> float32[n, 4]."""
    lut = np.ascontiguousarray(sky_view, dtype=np.float32)
    rec = np.ascontiguousarray(lighting16, dtype=np.float32)
    d = np.zeros((len(dirs), 4), dtype=np.float32)
    d[:, :3] = np.asarray(dirs, dtype=np.float32)[:, :3]
    out = np.zeros((len(d), 4), dtype=np.float32)
    liboracle().orc_sky_view_lookup(_p(lut, f32p), _p(rec, f32p), _p(d, f32p), len(d), _p(out, f32p))
    return out
