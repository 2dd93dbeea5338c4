"""CPU oracle for the texture-space path tracer.

TEST INFRASTRUCTURE ONLY.  Tests, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package as the checker;
the product package ``nan_amd`` never imports it.

``liboracle()`` loads ``oracle/libnan_oracle.so`` (built by ``make -C
oracle``) and declares its C signatures; ``OracleScene`` wraps one scene.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
f32p = ctypes.POINTER(ctypes.c_float)
f64p = ctypes.POINTER(ctypes.c_double)


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("vertices", f32p), ("vertex_count", ctypes.c_uint32),
        ("indices", u32p), ("index_count", ctypes.c_uint32),
        ("mesh_descs", u32p), ("mesh_count", ctypes.c_uint32),
        ("instance_descs", u32p), ("instance_count", ctypes.c_uint32),
        ("transforms", f32p), ("inv_transposes", f32p), ("transform_count", ctypes.c_uint32),
        ("materials", u32p), ("material_count", ctypes.c_uint32),
        ("lighting", f32p),
        ("textures", u32p), ("texture_count", ctypes.c_uint32),
        ("texels", u32p), ("texel_count", ctypes.c_uint64),
    ]


def build() -> Path:
    import subprocess
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _HERE / "libnan_oracle.so"


def liboracle():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = _HERE / "libnan_oracle.so"
    if not path.exists():
        build()
    lib = ctypes.CDLL(str(path))
    lib.orc_rng_stream.restype = ctypes.c_uint32
    lib.orc_rng_stream.argtypes = [ctypes.c_uint32] * 4 + [f32p]
    lib.orc_address_to_ij.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p]
    lib.orc_nearest_texel.restype = ctypes.c_uint32
    lib.orc_nearest_texel.argtypes = [f32p, u32p, u32p]
    lib.orc_sample_irradiance.argtypes = [f32p, u32p, f32p, u32p, u32p, u32p, ctypes.c_int, f32p]
    lib.orc_texel_raygen.restype = ctypes.c_int
    lib.orc_texel_raygen.argtypes = [u32p, u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p]
    lib.orc_pack_rgb9e5.restype = ctypes.c_uint32
    lib.orc_pack_rgb9e5.argtypes = [f32p]
    lib.orc_unpack_rgb9e5.argtypes = [ctypes.c_uint32, f32p]
    lib.orc_pack_rgb9e5_red_range.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p]
    lib.orc_pack_rgb9e5_array.argtypes = [f32p, ctypes.c_uint32, ctypes.c_uint64, u32p]
    lib.orc_scene_create.restype = ctypes.c_void_p
    lib.orc_scene_create.argtypes = [ctypes.POINTER(SceneDesc), ctypes.c_int]
    lib.orc_scene_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_scene_triangle_count.restype = ctypes.c_uint32
    lib.orc_scene_triangle_count.argtypes = [ctypes.c_void_p]
    lib.orc_scene_world_triangles.argtypes = [ctypes.c_void_p, f32p]
    lib.orc_intersect.argtypes = [ctypes.c_void_p, f32p, ctypes.c_uint64, f32p, f32p, u32p, u32p, u32p]
    lib.orc_intersect_fp64.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint64, f64p, f64p, u32p, u32p, f64p, f64p]
    lib.orc_trace_texels.restype = ctypes.c_int
    lib.orc_trace_texels.argtypes = [ctypes.c_void_p, u32p, u32p, u32p, ctypes.c_uint32, f32p,
                                     ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     u64p, ctypes.c_uint64, ctypes.c_int, u64p]
    lib.orc_resolve.argtypes = [ctypes.c_void_p, u32p, u32p, u32p, f32p, u32p, f32p, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, f32p]
    lib.orc_accumulate.argtypes = [f32p, f32p, f32p, ctypes.c_uint64, ctypes.c_int]
    lib.orc_tonemap.argtypes = [f32p, f32p, ctypes.c_uint64, ctypes.c_float]
    lib.orc_sincos_2pi.argtypes = [f32p, ctypes.c_uint64, f32p, f32p]
    lib.orc_set_crosscheck.argtypes = [ctypes.c_int]
    lib.orc_crosscheck_result.restype = ctypes.c_int
    lib.orc_crosscheck_result.argtypes = [f32p]
    lib.orc_scene_set_sky_view_lut.argtypes = [ctypes.c_void_p, f32p]
    lib.orc_scene_set_russian_roulette.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_scene_set_frame_index.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    lib.orc_scene_set_rng_mode.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    lib.orc_pcg_stream.restype = ctypes.c_uint32
    lib.orc_pcg_stream.argtypes = [ctypes.c_uint32] * 4 + [u32p]
    lib.orc_sample_texture.argtypes = [ctypes.c_void_p, ctypes.c_uint32, f32p, ctypes.c_uint64, f32p]
    lib.orc_alpha_hash.argtypes = [f32p, ctypes.c_uint64, ctypes.c_uint32, u32p]
    lib.orc_texel_lut.argtypes = [f32p]
    lib.orc_murmur_key_mix.restype = ctypes.c_uint32
    lib.orc_murmur_key_mix.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.orc_sky_view_lookup.argtypes = [f32p, f32p, f32p, ctypes.c_uint64, f32p]
    lib.orc_atmo_scattering.argtypes = [f32p, f32p]
    lib.orc_atmo_sphere_hit.restype = ctypes.c_float
    lib.orc_atmo_sphere_hit.argtypes = [f32p, f32p, ctypes.c_float]
    lib.orc_atmo_sun_transmittance.argtypes = [f32p, f32p, f32p]
    lib.orc_atmo_transmittance_lut.argtypes = [f32p]
    lib.orc_atmo_multiscattering_lut.argtypes = [f32p, f32p]
    lib.orc_atmo_multiscattering_texel.argtypes = [f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                   f32p]
    lib.orc_atmo_sky_view_lut.argtypes = [f32p, f32p, f32p, f32p]
    lib.orc_atmo_view_transmittance.argtypes = [f32p, f32p, f32p]
    _LIB = lib
    return lib


def _p(arr, ptype):
    if arr is None:
        return None
    return arr.ctypes.data_as(ptype)


def rng_stream(x: int, y: int, s: int, count: int) -> tuple[np.ndarray, int]:
    out = np.zeros(count, dtype=np.float32)
    state = liboracle().orc_rng_stream(x, y, s, count, _p(out, f32p))
    return out, int(state)


def pack_rgb9e5(values: np.ndarray) -> np.ndarray:
    values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 3)
    padded = np.zeros((len(values), 4), dtype=np.float32)
    padded[:, :3] = values
    out = np.zeros(len(values), dtype=np.uint32)
    liboracle().orc_pack_rgb9e5_array(_p(padded, f32p), 4, len(values), _p(out, u32p))
    return out


def pack_rgb9e5_red_range(start: int, count: int) -> np.ndarray:
    """Words for (x, 0, 0), x over float32 bit patterns start..start+count-1."""
    out = np.zeros(count, dtype=np.uint32)
    liboracle().orc_pack_rgb9e5_red_range(start, count, _p(out, u32p))
    return out


def unpack_rgb9e5(packed: int) -> np.ndarray:
    out = np.zeros(3, dtype=np.float32)
    liboracle().orc_unpack_rgb9e5(int(packed), _p(out, f32p))
    return out


def address_to_ij(address: int, resolution: int) -> tuple[int, int]:
    ij = np.zeros(2, dtype=np.uint32)
    liboracle().orc_address_to_ij(address, resolution, _p(ij, u32p))
    return int(ij[0]), int(ij[1])


def nearest_texel(bary, inst_row, face_row) -> int:
    """mesh_colors_nearest_texel for one barycentric triple (float32[3])
    on the face whose layout rows are given (instance uint32[4], face
    uint32[2]); returns the slot index."""
    b = np.ascontiguousarray(bary, dtype=np.float32).reshape(3)
    i = np.ascontiguousarray(inst_row, dtype=np.uint32).reshape(4)
    f = np.ascontiguousarray(face_row, dtype=np.uint32).reshape(2)
    return int(liboracle().orc_nearest_texel(_p(b, f32p), _p(i, u32p), _p(f, u32p)))


def sample_irradiance(bary, inst_row, face_row, side_row, back: bool = False, payload=None, packed=None):
    """The resolve's 3-tap lookup (mesh_colors_sample_irradiance) for one
    barycentric triple on one face, reading payload (float32[N,4]) or packed
    RGB9E5 words; returns float32[3]."""
    b = np.ascontiguousarray(bary, dtype=np.float32).reshape(3)
    i = np.ascontiguousarray(inst_row, dtype=np.uint32).reshape(4)
    f = np.ascontiguousarray(face_row, dtype=np.uint32).reshape(2)
    sd = np.ascontiguousarray(side_row, dtype=np.uint32).reshape(2)
    out = np.zeros(3, dtype=np.float32)
    pl = None if payload is None else np.ascontiguousarray(payload, dtype=np.float32)
    pk = None if packed is None else np.ascontiguousarray(packed, dtype=np.uint32)
    liboracle().orc_sample_irradiance(_p(pl, f32p), _p(pk, u32p), _p(b, f32p), _p(i, u32p), _p(f, u32p),
                                      _p(sd, u32p), 1 if back else 0, _p(out, f32p))
    return out


def texel_raygen(faces, instances, iteration: int, spp: int):
    faces = np.ascontiguousarray(faces, dtype=np.uint32)
    instances = np.ascontiguousarray(instances, dtype=np.uint32)
    total = int(instances[:, 3].astype(np.int64).sum())
    out = np.zeros((total, 4), dtype=np.uint32)
    rng = np.zeros(total, dtype=np.uint32)
    rc = liboracle().orc_texel_raygen(_p(faces, u32p), _p(instances, u32p), len(instances), iteration, spp,
                                      _p(out, u32p), _p(rng, u32p))
    if rc != 0:
        raise RuntimeError("oracle raygen failed")
    return out, rng


def sincos_2pi(u: np.ndarray):
    u = np.ascontiguousarray(u, dtype=np.float32)
    s = np.zeros_like(u)
    c = np.zeros_like(u)
    liboracle().orc_sincos_2pi(_p(u, f32p), len(u), _p(s, f32p), _p(c, f32p))
    return s, c


def pcg_stream(x: int, y: int, s: int, count: int) -> tuple[np.ndarray, int]:
    """The first ``count`` raw 32-bit draws of the PCG stream of texel x,
    instance y and sample s, and the state after them."""
    out = np.zeros(count, dtype=np.uint32)
    state = liboracle().orc_pcg_stream(x & 0xFFFFFFFF, y & 0xFFFFFFFF, s & 0xFFFFFFFF, count, _p(out, u32p))
    return out, int(state)


def texel_lut() -> np.ndarray:
    """The oracle's texel decode table float32[512]: c / 255, then the sRGB
    decode of c / 255."""
    out = np.zeros(512, dtype=np.float32)
    liboracle().orc_texel_lut(_p(out, f32p))
    return out


def alpha_hash(positions: np.ndarray, frame_index: int) -> np.ndarray:
    """murmur_key_mix(position_hashing(p, 10000), frame_index) per
    position float32[n, 3] (common.slang:540-577)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    out = np.zeros(len(pos), dtype=np.uint32)
    liboracle().orc_alpha_hash(_p(pos, f32p), len(pos), int(frame_index) & 0xFFFFFFFF, _p(out, u32p))
    return out


class OracleScene:
    """One scene in the oracle: world-space triangles plus an optional BVH.

    ``arrays`` is the flat dictionary produced by
    ``nan_amd.scene.flatten_scene`` (vertices, indices, mesh_descs,
    instance_descs, transforms, inv_transposes, materials, lighting)."""

    def __init__(self, arrays: dict, use_bvh: bool = True):
        self.lib = liboracle()
        self.arrays = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        a = self.arrays
        self.desc = SceneDesc(
            _p(a["vertices"], f32p), a["vertices"].shape[0],
            _p(a["indices"], u32p), a["indices"].size,
            _p(a["mesh_descs"], u32p), a["mesh_descs"].shape[0],
            _p(a["instance_descs"], u32p), a["instance_descs"].shape[0],
            _p(a["transforms"], f32p), _p(a["inv_transposes"], f32p), a["transforms"].shape[0],
            _p(a["materials"], u32p), a["materials"].shape[0],
            _p(a["lighting"], f32p),
            _p(a["textures"], u32p), a["textures"].shape[0],
            _p(a["texels"], u32p), a["texels"].size,
        )
        self.handle = self.lib.orc_scene_create(ctypes.byref(self.desc), 1 if use_bvh else 0)
        if bool(arrays.get("russian_roulette", False)):
            self.set_russian_roulette(True)

    def set_russian_roulette(self, enable: bool) -> None:
        """Russian roulette on the scene's paths (the rule of trace_radiance)."""
        self.lib.orc_scene_set_russian_roulette(self.handle, 1 if enable else 0)

    def set_rng(self, rng: str) -> None:
        """Random streams of the texel paths: "reference" (the LCG) or "pcg"."""
        self.lib.orc_scene_set_rng_mode(self.handle, {"reference": 0, "pcg": 1}[rng])

    def set_frame_index(self, frame_index: int) -> None:
        """Frame index of the blend alpha test (scene.slang:146-148)."""
        self.lib.orc_scene_set_frame_index(self.handle, int(frame_index) & 0xFFFFFFFF)

    def sample_texture(self, texture: int, uv: np.ndarray) -> np.ndarray:
        """Bilinear, wrap-addressed samples float32[n, 4] of the scene's
        texture ``texture`` at ``uv`` float32[n, 2]."""
        uv = np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 2)
        out = np.zeros((len(uv), 4), dtype=np.float32)
        self.lib.orc_sample_texture(self.handle, int(texture), _p(uv, f32p), len(uv), _p(out, f32p))
        return out

    def set_sky_view_lut(self, lut) -> None:
        """Attaches a sky-view table float32[200, 200, 4] (None detaches)."""
        self.sky_lut = None if lut is None else np.ascontiguousarray(lut, dtype=np.float32).reshape(200, 200, 4)
        self.lib.orc_scene_set_sky_view_lut(self.handle, _p(self.sky_lut, f32p))

    def close(self):
        if self.handle:
            self.lib.orc_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def triangle_count(self) -> int:
        return int(self.lib.orc_scene_triangle_count(self.handle))

    def world_triangles(self) -> np.ndarray:
        out = np.zeros((self.triangle_count, 9), dtype=np.float32)
        self.lib.orc_scene_world_triangles(self.handle, _p(out, f32p))
        return out

    def intersect(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        t = np.zeros(n, dtype=np.float32)
        b = np.zeros((n, 2), dtype=np.float32)
        inst = np.zeros(n, dtype=np.uint32)
        prim = np.zeros(n, dtype=np.uint32)
        occ = np.zeros(n, dtype=np.uint32)
        self.lib.orc_intersect(self.handle, _p(rays, f32p), n, _p(t, f32p), _p(b, f32p), _p(inst, u32p),
                               _p(prim, u32p), _p(occ, u32p))
        return t, b, inst, prim, occ

    def intersect_fp64(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float64).reshape(-1, 8)
        n = len(rays)
        t = np.zeros(n)
        b = np.zeros((n, 2))
        inst = np.zeros(n, dtype=np.uint32)
        prim = np.zeros(n, dtype=np.uint32)
        gap = np.zeros(n)
        edge = np.zeros(n)
        self.lib.orc_intersect_fp64(self.handle, _p(rays, f64p), n, _p(t, f64p), _p(b, f64p), _p(inst, u32p),
                                    _p(prim, u32p), _p(gap, f64p), _p(edge, f64p))
        return t, b, inst, prim, gap, edge

    def trace_texels(self, faces, instances, sides, payload, iteration, spp, max_bounces, reset,
                     work=None, threads=1):
        """Runs the plain texel tracer on ``payload`` (float32[N,4] viewed
        as slots) in place; returns [extend rays, shadow rays, tri tests,
        slot updates]."""
        faces = np.ascontiguousarray(faces, dtype=np.uint32)
        instances = np.ascontiguousarray(instances, dtype=np.uint32)
        sides = np.ascontiguousarray(sides, dtype=np.uint32)
        assert payload.dtype == np.float32 and payload.flags.c_contiguous
        counters = np.zeros(4, dtype=np.uint64)
        wk = None if work is None else np.ascontiguousarray(work, dtype=np.uint64)
        rc = self.lib.orc_trace_texels(self.handle, _p(faces, u32p), _p(instances, u32p), _p(sides, u32p),
                                       len(instances), _p(payload, f32p), iteration, spp, max_bounces,
                                       1 if reset else 0, _p(wk, u64p), 0 if wk is None else len(wk),
                                       int(threads), _p(counters, u64p))
        if rc != 0:
            raise RuntimeError("oracle tracer failed")
        return counters

    def resolve(self, faces, instances, sides, camera, width, height, payload=None, packed=None,
                rows=None):
        faces = np.ascontiguousarray(faces, dtype=np.uint32)
        instances = np.ascontiguousarray(instances, dtype=np.uint32)
        sides = np.ascontiguousarray(sides, dtype=np.uint32)
        camera = np.ascontiguousarray(camera, dtype=np.float32)
        out = np.zeros((height, width, 4), dtype=np.float32)
        r0, r1 = (0, height) if rows is None else rows
        self.lib.orc_resolve(self.handle, _p(faces, u32p), _p(instances, u32p), _p(sides, u32p),
                             _p(payload, f32p), _p(packed, u32p), _p(camera, f32p), width, height, r0, r1,
                             _p(out, f32p))
        return out


def accumulate(inp: np.ndarray, accum: np.ndarray, reset: bool) -> np.ndarray:
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    out = np.zeros_like(inp)
    liboracle().orc_accumulate(_p(inp, f32p), _p(accum, f32p), _p(out, f32p), inp.size // 4, 1 if reset else 0)
    return out


def tonemap(inp: np.ndarray, exposure: float) -> np.ndarray:
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    out = np.zeros_like(inp)
    liboracle().orc_tonemap(_p(inp, f32p), _p(out, f32p), inp.size // 4, exposure)
    return out


def cpu_count() -> int:
    """Worker threads for the oracle: the process's CPUs, at most 16 (the
    host share of one GPU on a shared machine)."""
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))
