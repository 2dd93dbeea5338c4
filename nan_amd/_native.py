"""ctypes binding of ``libnan_hip.so`` (declared in ``include/nan_hip.h``).

Loading follows the reference's native helpers (surface_probe_sampler.py:
251-350): a CDLL with typed signatures and a typed Python exception carrying
the library's error text.  There is no fallback: if the library is missing
or cannot be loaded, :func:`lib` raises :class:`NativeLibraryError`.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libnan_hip.so"
ABI_VERSION = 3  # NAN_HIP_ABI_VERSION of include/nan_hip.h

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
f32p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p


class NativeLibraryError(RuntimeError):
    """libnan_hip.so is missing or failed to load."""


class NanError(RuntimeError):
    """A libnan_hip call returned a non-zero status."""


class NanSceneDesc(ctypes.Structure):
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


class NanBvhOptions(ctypes.Structure):
    _fields_ = [("sah_bins", ctypes.c_uint32), ("max_leaf_size", ctypes.c_uint32),
                ("build_threads", ctypes.c_uint32)]


class NanSceneInfo(ctypes.Structure):
    _fields_ = [
        ("triangle_count", ctypes.c_uint64), ("node_count", ctypes.c_uint64), ("leaf_count", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64), ("max_stack", ctypes.c_uint32), ("stack_depth", ctypes.c_uint32),
        ("build_ms", ctypes.c_double), ("sah_cost", ctypes.c_float),
    ]


class NanNrcConfig(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_uint32), ("features_per_level", ctypes.c_uint32), ("log2_table_size", ctypes.c_uint32),
        ("base_resolution", ctypes.c_float), ("max_resolution", ctypes.c_float),
        ("bbox_min", ctypes.c_float * 3), ("bbox_max", ctypes.c_float * 3),
        ("width", ctypes.c_uint32), ("hidden_layers", ctypes.c_uint32),
        ("learning_rate", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
        ("epsilon", ctypes.c_float), ("seed", ctypes.c_uint32), ("dense_table_adam", ctypes.c_uint32),
        ("minibatches", ctypes.c_uint32),
    ]


COUNTER_EXTEND_RAYS = 0
COUNTER_SHADOW_RAYS = 1
COUNTER_NODE_VISITS = 2
COUNTER_TRI_TESTS = 3
COUNTER_SLOT_UPDATES = 4
COUNTER_SHADOW_NODE_VISITS = 5
COUNTER_SHADOW_TRI_TESTS = 6
COUNTER_COUNT = 8

STAGES = ("generate", "extend", "shade", "shadow", "finish", "megakernel", "learned_paths", "cache_train",
          "first_bounce", "cache_query")

LEARNED_EMITTER_NEE = 1
LEARNED_RAMPED_MEAN = 2
LEARNED_TEXEL_QUERY = 4
LEARNED_RNG_PCG = 8
LEARNED_OVERLAP = 16
RNG_MODES = {"reference": 0, "pcg": 1}  # NAN_RNG_REFERENCE, NAN_RNG_PCG

TRACER_MEGAKERNEL = 0
TRACER_WAVEFRONT = 1

# every symbol include/nan_hip.h declares
EXPORTED_SYMBOLS = (
    "nan_version", "nan_last_error", "nan_scene_create", "nan_scene_destroy", "nan_scene_info",
    "nan_scene_set_lighting", "nan_trace_rays", "nan_tracer_create", "nan_tracer_destroy", "nan_tracer_tables",
    "nan_tracer_set_mode", "nan_tracer_profile", "nan_tracer_stage_times", "nan_tst_execute", "nan_tst_execute_counted", "nan_tst_raygen", "nan_rgb9e5_pack", "nan_rgb9e5_unpack",
    "nan_resolve", "nan_accumulate", "nan_tonemap", "nan_tracer_tune", "nan_batch_create",
    "nan_batch_set_payloads", "nan_batch_set_bounces", "nan_tracer_device_frame",
    "nan_nrc_create", "nan_nrc_destroy", "nan_nrc_param_counts", "nan_nrc_get_params", "nan_nrc_set_params",
    "nan_nrc_infer", "nan_nrc_infer_noalloc", "nan_nrc_train", "nan_nrc_forward_raw", "nan_nrc_backward_raw", "nan_nrc_encode",
    "nan_nrc_stats", "nan_nrc_time_forward", "nan_tst_execute_learned", "nan_surface_filter",
    "nan_scene_set_sky_view_lut", "nan_atmosphere_transmittance_lut", "nan_atmosphere_multiscattering_lut",
    "nan_atmosphere_sky_view_lut", "nan_atmosphere_view_transmittance", "nan_sky_view_lookup",
    "nan_tst_learned_records", "nan_tst_learned_reserve", "nan_scene_set_russian_roulette", "nan_calibrate_valu",
    "nan_scene_set_frame_index", "nan_batch_set_frame_index", "nan_nrc_table_grad", "nan_nrc_set_fused_training",
)

_LIB = None


def _declare(lib) -> None:
    I = ctypes.c_int
    U32 = ctypes.c_uint32
    U64 = ctypes.c_uint64
    sig = {
        "nan_version": (I, []),
        "nan_last_error": (ctypes.c_char_p, []),
        "nan_scene_create": (I, [ctypes.POINTER(NanSceneDesc), ctypes.POINTER(NanBvhOptions), ctypes.POINTER(vp),
                                 ctypes.c_char_p, ctypes.c_size_t]),
        "nan_scene_destroy": (None, [vp]),
        "nan_scene_info": (I, [vp, ctypes.POINTER(NanSceneInfo)]),
        "nan_scene_set_lighting": (I, [vp, f32p]),
        "nan_scene_set_sky_view_lut": (I, [vp, vp]),
        "nan_scene_set_russian_roulette": (I, [vp, ctypes.c_uint32]),
        "nan_scene_set_frame_index": (I, [vp, ctypes.c_uint32, vp]),
        "nan_nrc_table_grad": (I, [vp, vp, vp, vp, vp, U32, vp, vp]),
        "nan_batch_set_frame_index": (I, [vp, ctypes.c_uint32, vp]),
        "nan_atmosphere_transmittance_lut": (I, [vp, vp]),
        "nan_atmosphere_multiscattering_lut": (I, [vp, vp, vp]),
        "nan_atmosphere_sky_view_lut": (I, [vp, vp, f32p, vp, vp]),
        "nan_atmosphere_view_transmittance": (I, [vp, f32p, vp, vp]),
        "nan_sky_view_lookup": (I, [vp, f32p, vp, U64, vp, vp]),
        "nan_tst_learned_records": (I, [vp, U32, U32, U32, U32, vp, vp, vp, vp, vp, vp]),
        "nan_trace_rays": (I, [vp, vp, U64, vp, vp, vp, vp]),
        "nan_tracer_create": (I, [vp, u32p, U32, u32p, u32p, U32, U64, U64, ctypes.POINTER(vp), ctypes.c_char_p,
                                  ctypes.c_size_t]),
        "nan_tracer_destroy": (None, [vp]),
        "nan_tracer_tables": (I, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]),
        "nan_tracer_set_mode": (I, [vp, I]),
        "nan_tracer_profile": (I, [vp, I]),
        "nan_tracer_stage_times": (I, [vp, ctypes.POINTER(ctypes.c_double), u32p]),
        "nan_tst_execute": (I, [vp, vp, U32, U32, U32, U32, U32, vp]),
        "nan_tst_execute_counted": (I, [vp, vp, U32, U32, U32, U32, U32, vp, vp]),
        "nan_tst_raygen": (I, [vp, U32, U32, vp, vp, vp]),
        "nan_rgb9e5_pack": (I, [vp, vp, U64, vp]),
        "nan_rgb9e5_unpack": (I, [vp, vp, U64, vp]),
        "nan_resolve": (I, [vp, vp, I, f32p, vp, U32, U32, vp]),
        "nan_accumulate": (I, [vp, vp, vp, U64, I, vp]),
        "nan_tonemap": (I, [vp, vp, U64, ctypes.c_float, vp]),
        "nan_tracer_tune": (I, [vp, U32, U32, U32]),
        "nan_batch_create": (I, [ctypes.POINTER(vp), ctypes.POINTER(vp), U32, ctypes.POINTER(vp), ctypes.c_char_p,
                                 ctypes.c_size_t]),
        "nan_batch_set_payloads": (I, [vp, ctypes.POINTER(vp)]),
        "nan_batch_set_bounces": (I, [vp, u32p]),
        "nan_tracer_device_frame": (I, [vp, I, U32, U32]),
        "nan_nrc_create": (I, [ctypes.POINTER(NanNrcConfig), ctypes.POINTER(vp), ctypes.c_char_p, ctypes.c_size_t]),
        "nan_nrc_destroy": (None, [vp]),
        "nan_nrc_param_counts": (I, [vp, u64p, u64p]),
        "nan_nrc_get_params": (I, [vp, vp, vp]),
        "nan_nrc_set_params": (I, [vp, vp, vp]),
        "nan_nrc_infer": (I, [vp, vp, vp, vp, vp, U32, vp]),
        "nan_nrc_infer_noalloc": (I, [vp, vp, vp, vp, vp, U32, vp]),
        "nan_calibrate_valu": (I, [U32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
        "nan_nrc_train": (I, [vp, vp, vp, vp, vp, U32, vp, vp]),
        "nan_nrc_forward_raw": (I, [vp, vp, vp, vp, vp, U32, vp]),
        "nan_nrc_backward_raw": (I, [vp, vp, vp, vp, vp, U32, vp, vp, vp]),
        "nan_nrc_encode": (I, [vp, vp, vp, vp, vp, U32, vp]),
        "nan_nrc_stats": (I, [vp, f32p]),
        "nan_nrc_set_fused_training": (I, [vp, I]),
        "nan_nrc_time_forward": (I, [vp, vp, vp, vp, U32, U32, f32p, vp]),
        "nan_tst_execute_learned": (I, [vp, vp, vp, U32, U32, U32, U32, U32, U32, vp, vp]),
        "nan_tst_learned_reserve": (I, [vp, vp, U32, U32]),
        "nan_surface_filter": (I, [vp, vp, vp, vp, vp, ctypes.c_float, ctypes.c_float, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """Loads libnan_hip.so once; raises NativeLibraryError if it is absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise NativeLibraryError(
            f"{LIB_PATH} is not built; run `python -m nan_amd.build` (hipcc, gfx950)")
    try:
        handle = ctypes.CDLL(str(LIB_PATH))
    except OSError as exc:
        raise NativeLibraryError(f"cannot load {LIB_PATH}: {exc}") from exc
    _declare(handle)
    if handle.nan_version() != ABI_VERSION:
        raise NativeLibraryError(f"{LIB_PATH} has ABI version {handle.nan_version()}, this package needs "
                                 f"{ABI_VERSION}; rebuild it (`python -m nan_amd.build`)")
    _LIB = handle
    return handle


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().nan_last_error()
        raise NanError(f"{what}: {msg.decode() if msg else 'unknown error'}")


def err_buffer():
    return ctypes.create_string_buffer(512)
