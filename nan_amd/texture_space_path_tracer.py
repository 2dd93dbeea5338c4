"""Texture-space path tracer pass (texture_space_path_tracer.py:7-59 of the
reference), backed by the gfx950 kernels of libnan_hip.

The reference dispatches ``compute_main`` once per instance
(texture_space_path_tracer.py:42-59).  Here one launch covers every front
texel of every instance through a flat texel->face map; the per-texel work
is the reference's (texture_space_path_tracer.slang:140-216) and every slot
update is bit-identical to the CPU oracle.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .device import CommandEncoder, Device
from .mesh_colors import MESH_COLORS_PAYLOAD_SIZE, MeshColorsLayout
from .scene import Scene


def _u32(a):
    return a.ctypes.data_as(_native.u32p)


def as_payload_tensor(buf: torch.Tensor, slots: int) -> torch.Tensor:
    """Checks a 16-byte-per-slot payload buffer (float32[N,4] or any
    contiguous tensor of ``16 * N`` bytes) on the device."""
    if not buf.is_cuda or not buf.is_contiguous():
        raise ValueError("payload must be a contiguous device tensor")
    if buf.numel() * buf.element_size() < slots * MESH_COLORS_PAYLOAD_SIZE:
        raise ValueError(f"payload holds fewer than {slots} 16-byte slots")
    return buf


class TextureSpacePathTracer:
    """Progressive per-texel irradiance tracer (texture_space_path_tracer.py:7-59)."""

    MODES = {"megakernel": _native.TRACER_MEGAKERNEL, "wavefront": _native.TRACER_WAVEFRONT}

    def __init__(self, device: Device, scene: Scene, layout: MeshColorsLayout, *, samples_per_texel: int = 1,
                 max_bounces: int = 3, mode: str = "wavefront", rng: str = "reference"):
        self.device = device
        self.scene = scene
        self.layout = layout
        self.samples_per_texel = max(1, int(samples_per_texel))
        self.max_bounces = max(1, int(max_bounces))
        self.set_rng(rng)
        handle = ctypes.c_void_p()
        err = _native.err_buffer()
        rc = _native.lib().nan_tracer_create(
            scene.handle, _u32(layout.face_table), len(layout.face_table), _u32(layout.instance_table),
            _u32(layout.side_table), len(layout.instance_table), layout.total_surface_texels,
            layout.total_payload_count, ctypes.byref(handle), err, len(err))
        if rc != 0:
            raise _native.NanError(err.value.decode())
        self.handle = handle
        self.set_mode(mode)
        dev = device.torch_device
        # The reference exposes its face/instance/side/adjacency buffers for the
        # resolve and filter passes (texture_space_path_tracer.py:26-31).
        self.face_infos_buffer = torch.from_numpy(layout.face_table.view(np.int32)).to(dev)
        self.instance_infos_buffer = torch.from_numpy(layout.instance_table.view(np.int32)).to(dev)
        self.side_infos_buffer = torch.from_numpy(layout.side_table.view(np.int32)).to(dev)
        self.adjacency_infos_buffer = torch.from_numpy(layout.adjacency_table.view(np.int32)).to(dev)

    def set_rng(self, rng: str) -> None:
        """Random streams of the texel paths: ``"reference"`` (the
        reference's LCG seeded per texel, instance and sample,
        common.slang:164-210) or ``"pcg"`` (PCG, RXS-M-XS output, started at
        a hash of texel, instance and sample; no reference counterpart)."""
        if rng not in _native.RNG_MODES:
            raise ValueError(f"unknown rng {rng!r}")
        self.rng = rng

    def set_mode(self, mode: str) -> None:
        """``"wavefront"`` (staged kernels, default) or ``"megakernel"`` (one
        thread per texel); both produce bit-identical payloads."""
        if mode not in self.MODES:
            raise ValueError(f"unknown tracer mode {mode!r}")
        _native.check(_native.lib().nan_tracer_set_mode(self.handle, self.MODES[mode]), "nan_tracer_set_mode")
        self.mode = mode

    def close(self) -> None:
        if getattr(self, "handle", None):
            _native.lib().nan_tracer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, command_encoder: CommandEncoder, output: torch.Tensor, iteration: int, *,
                reset: bool = False, counters: torch.Tensor | None = None) -> None:
        """One iteration over all texels into ``output`` (16-byte slots)."""
        as_payload_tensor(output, self.layout.total_payload_count)
        stream = command_encoder.stream_handle
        it = max(0, int(iteration))
        if counters is None:
            rc = _native.lib().nan_tst_execute(self.handle, output.data_ptr(), it, self.samples_per_texel,
                                               self.max_bounces, 1 if reset else 0, _native.RNG_MODES[self.rng],
                                               stream)
        else:
            if counters.dtype != torch.int64 or counters.numel() < _native.COUNTER_COUNT or not counters.is_cuda:
                raise ValueError("counters must be an int64 device tensor of at least 8 entries")
            rc = _native.lib().nan_tst_execute_counted(self.handle, output.data_ptr(), it, self.samples_per_texel,
                                                       self.max_bounces, 1 if reset else 0,
                                                       _native.RNG_MODES[self.rng], counters.data_ptr(), stream)
        _native.check(rc, "nan_tst_execute")

    def learned_training_records(self, iteration: int, n: int, *, emitter_nee: bool = True,
                                 texel_query: bool = False) -> dict:
        """The ``n`` training records ``execute_learned`` would build at
        ``iteration`` (features ``pos``/``nrm``/``alb`` and ``target`` float32
        [n, 3], ``valid`` int32[n]).  Test hook; no counterpart in the
        reference."""
        dev = self.device.torch_device
        out = {k: torch.empty((n, 3), dtype=torch.float32, device=dev) for k in ("pos", "nrm", "alb", "target")}
        out["valid"] = torch.empty((n,), dtype=torch.int32, device=dev)
        flags = (_native.LEARNED_EMITTER_NEE if emitter_nee else 0) | (_native.LEARNED_TEXEL_QUERY if texel_query else 0)
        flags |= _native.LEARNED_RNG_PCG if self.rng == "pcg" else 0
        rc = _native.lib().nan_tst_learned_records(
            self.handle, max(0, int(iteration)), int(n), self.max_bounces, flags, out["pos"].data_ptr(),
            out["nrm"].data_ptr(), out["alb"].data_ptr(), out["target"].data_ptr(), out["valid"].data_ptr(),
            self.device.stream_handle)
        _native.check(rc, "nan_tst_learned_records")
        return out

    def execute_learned(self, command_encoder: CommandEncoder, output: torch.Tensor, iteration: int, cache, *,
                        reset: bool = False, train_samples: int = 65536, emitter_nee: bool = True,
                        ramped_mean: bool = True, texel_query: bool = False, overlap: bool = False,
                        counters: torch.Tensor | None = None) -> None:
        """One iteration of the learned form: trains ``cache`` (a
        :class:`~nan_amd.radiance_cache.RadianceCache`) on ``train_samples``
        plain path suffixes, then every texel sample takes the sun term and
        one bounce and ends in the cache.  ``emitter_nee`` adds light
        sampling of emissive triangles with MIS (same expectation, lower
        variance where emitters light the scene); ``ramped_mean`` weights the
        k-th iteration since the reset by k in the slot's mean, so the early,
        less trained cache fades out.  ``texel_query`` ends paths in the
        cache at the texel itself: the cache learns the indirect irradiance
        field and each texel adds only its sun and light-sampled emitter
        terms (no bounce ray; use a cache trained in this form only).
        ``overlap`` (one-bounce form) copies the cache's parameters at the
        start of the call, then trains the cache on the previous call's path
        suffixes on a second stream while this call traces and queries the
        copy: the queries see the cache as the previous call's training left
        it (trained on suffixes up to two frames old) and never wait for the
        training.  Before the first such call the kept records are zeros
        (none valid), so its step leaves the parameters unchanged.  No
        counterpart in the reference."""
        as_payload_tensor(output, self.layout.total_payload_count)
        if counters is not None and (counters.dtype != torch.int64 or counters.numel() < _native.COUNTER_COUNT
                                     or not counters.is_cuda):
            raise ValueError("counters must be an int64 device tensor of at least 8 entries")
        if not torch.cuda.is_current_stream_capturing():
            self.reserve_learned(cache, train_samples)
        rc = _native.lib().nan_tst_execute_learned(
            self.handle, cache.handle, output.data_ptr(), max(0, int(iteration)), self.samples_per_texel,
            self.max_bounces, 1 if reset else 0, int(train_samples),
            (_native.LEARNED_EMITTER_NEE if emitter_nee else 0) | (_native.LEARNED_RAMPED_MEAN if ramped_mean else 0)
            | (_native.LEARNED_TEXEL_QUERY if texel_query else 0) | (_native.LEARNED_RNG_PCG if self.rng == "pcg" else 0)
            | (_native.LEARNED_OVERLAP if overlap else 0),
            None if counters is None else counters.data_ptr(), command_encoder.stream_handle)
        _native.check(rc, "nan_tst_execute_learned")

    def set_device_frame(self, iteration: int, reset: bool = False) -> None:
        """Device-frame mode: execute and execute_learned take the iteration
        and reset flag from a device counter (set here) and advance it at the
        end of each call, so one captured graph of a call replays successive
        iterations.  No counterpart in the reference."""
        _native.check(_native.lib().nan_tracer_device_frame(self.handle, 1, int(iteration), 1 if reset else 0),
                      "nan_tracer_device_frame")

    def clear_device_frame(self) -> None:
        _native.check(_native.lib().nan_tracer_device_frame(self.handle, 0, 0, 0), "nan_tracer_device_frame")

    def capture_learned(self, command_encoder: CommandEncoder, output: torch.Tensor, cache, next_iteration: int, *,
                        reset: bool = False, train_samples: int = 65536, **kw) -> "torch.cuda.CUDAGraph":
        """Captures one learned iteration (``execute_learned`` with ``kw``)
        into a hipGraph in device-frame mode: replays run iterations
        next_iteration, next_iteration + 1, ..., the first resetting the
        slots when ``reset``, each training the cache once more (Adam's step
        count lives on the device).  The working set is reserved first, so
        this may be the tracer's first learned iteration.  No counterpart in
        the reference."""
        self.reserve_learned(cache, train_samples)
        self.set_device_frame(next_iteration, reset)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.execute_learned(command_encoder, output, 0, cache, train_samples=train_samples, **kw)
        return graph

    def reserve_learned(self, cache, train_samples: int = 65536) -> None:
        """Sizes the learned form's buffers (training records, cache queries,
        network scratch) for ``train_samples`` training paths per iteration.
        ``execute_learned`` does this itself outside graph capture; call it
        before capturing a first learned iteration into a graph, since a
        captured call allocates nothing and refuses to grow them.  No
        counterpart in the reference."""
        _native.check(_native.lib().nan_tst_learned_reserve(self.handle, cache.handle, self.samples_per_texel,
                                                            int(train_samples)), "nan_tst_learned_reserve")

    def tune(self, blocks_per_cu: int = 0, lds_stack: int = 16, refill_steps: int = 16) -> None:
        """Launch shape of the traversal kernels (results do not depend on it)."""
        _native.check(_native.lib().nan_tracer_tune(self.handle, int(blocks_per_cu), int(lds_stack),
                                                    int(refill_steps)), "nan_tracer_tune")
        self.tuning = (int(blocks_per_cu), int(lds_stack), int(refill_steps))

    def profile(self, enable: bool) -> None:
        """Bracket every kernel launch with HIP events (see stage_times)."""
        _native.check(_native.lib().nan_tracer_profile(self.handle, 1 if enable else 0), "nan_tracer_profile")

    def stage_times(self) -> dict:
        """Milliseconds and launches per stage since profiling was enabled
        (waits for the recorded events)."""
        ms = (ctypes.c_double * len(_native.STAGES))()
        n = (ctypes.c_uint32 * len(_native.STAGES))()
        _native.check(_native.lib().nan_tracer_stage_times(self.handle, ms, n), "nan_tracer_stage_times")
        return {name: (ms[k], n[k]) for k, name in enumerate(_native.STAGES) if n[k]}

    def raygen(self, command_encoder: CommandEncoder, iteration: int) -> tuple[torch.Tensor, torch.Tensor]:
        """Texel ray generation only: int32[front, 4] (instance, primitive,
        bary.x bits, bary.y bits) and the RNG state after jitter."""
        n = self.layout.total_surface_texels
        dev = self.device.torch_device
        out = torch.empty((n, 4), dtype=torch.int32, device=dev)
        rng = torch.empty((n,), dtype=torch.int32, device=dev)
        _native.check(_native.lib().nan_tst_raygen(self.handle, int(iteration), self.samples_per_texel,
                                                   out.data_ptr(), rng.data_ptr(), command_encoder.stream_handle),
                      "nan_tst_raygen")
        return out, rng
