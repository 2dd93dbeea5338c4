"""Learned radiance cache (hash-grid encoding + fused MFMA MLP) backed by the
``nan_nrc_*`` entry points of libnan_hip.

There is no counterpart in the reference renderer: the cache serves the
learned form of the texel tracer (:meth:`TextureSpacePathTracer.
execute_learned`), where every texel sample ends in the cache after one
bounce and plain path suffixes train it online.

Network: ``levels`` hash-grid levels of 2 features (``2**log2_table_size``
FP16 entries per level, resolutions growing geometrically from
``base_resolution`` to ``max_resolution`` over the scene box), concatenated
with normal and albedo, zero-padded to 48 inputs; four ReLU layers of 64 and
an RGB output (padded to 16), no biases; FP16 on MFMA with FP32
accumulation.  Loss: relative L2, ``sum_c (p - t)^2 / (p^2 + 0.01)`` with the
denominator held constant, averaged over records and channels.  Optimiser:
Adam on FP32 master parameters.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _native
from .device import CommandEncoder, Device

LAYER_SHAPES = ((64, 48), (64, 64), (64, 64), (64, 64), (16, 64))  # (out, in) per layer
INPUT_WIDTH = 48


@dataclass
class RadianceCacheConfig:
    levels: int = 16
    log2_table_size: int = 19
    base_resolution: float = 16.0
    max_resolution: float = 512.0
    learning_rate: float = 1e-2
    beta1: float = 0.9
    beta2: float = 0.99
    epsilon: float = 1e-15
    seed: int = 1337
    dense_table_adam: bool = False
    minibatches: int = 1

    @classmethod
    def for_texel_density(cls, extent: float, texels_per_unit: float, **kw) -> "RadianceCacheConfig":
        """Grid levels matched to a Mesh Colors layout: the finest level
        has about one cell per texel across the scene's longest extent
        (``extent * texels_per_unit``, clamped to [16, 512]) and the levels
        grow by about 2^(1/3) each, as the default 16 levels from 16 to 512
        do.  Every level costs 8 hash-table gradient atomics per training
        record, so a small scene trains several times faster with the few
        levels its texels can resolve (measured on the Cornell box: 4 levels
        to 32 cells, 0.41 against 0.63 ms per learned frame at the same RMSE
        after 256 frames)."""
        base = float(kw.get("base_resolution", 16.0))
        finest = float(np.clip(extent * texels_per_unit, base, 512.0))
        levels = 1 + int(round(3.0 * math.log2(finest / base)))
        return cls(levels=min(16, max(1, levels)), max_resolution=finest, **kw)


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def _check_rows(t: torch.Tensor, n: int, cols: int, name: str, device: torch.device) -> None:
    if t.dtype != torch.float32 or not t.is_contiguous() or tuple(t.shape) != (n, cols):
        raise ValueError(f"{name} must be a contiguous float32[{n}, {cols}] tensor")
    if t.device != device:
        raise ValueError(f"{name} must live on {device}")


class RadianceCache:
    """One learned radiance cache on a device."""

    def __init__(self, device: Device, bbox_min, bbox_max, config: RadianceCacheConfig | None = None):
        self.device = device
        self.config = config or RadianceCacheConfig()
        c = self.config
        cfg = _native.NanNrcConfig()
        cfg.levels = c.levels
        cfg.features_per_level = 2
        cfg.log2_table_size = c.log2_table_size
        cfg.base_resolution = c.base_resolution
        cfg.max_resolution = c.max_resolution
        for a in range(3):
            cfg.bbox_min[a] = float(bbox_min[a])
            cfg.bbox_max[a] = float(bbox_max[a])
        cfg.width = 64
        cfg.hidden_layers = 4
        cfg.learning_rate = c.learning_rate
        cfg.beta1 = c.beta1
        cfg.beta2 = c.beta2
        cfg.epsilon = c.epsilon
        cfg.seed = c.seed
        cfg.dense_table_adam = 1 if c.dense_table_adam else 0
        cfg.minibatches = c.minibatches
        handle = ctypes.c_void_p()
        err = _native.err_buffer()
        rc = _native.lib().nan_nrc_create(ctypes.byref(cfg), ctypes.byref(handle), err, len(err))
        if rc != 0:
            raise _native.NanError(err.value.decode())
        self.handle = handle
        self.bbox_min = np.asarray(bbox_min, dtype=np.float32)
        self.bbox_max = np.asarray(bbox_max, dtype=np.float32)
        w = ctypes.c_uint64()
        t = ctypes.c_uint64()
        _native.check(_native.lib().nan_nrc_param_counts(self.handle, ctypes.byref(w), ctypes.byref(t)),
                      "nan_nrc_param_counts")
        self.weight_count = int(w.value)
        self.table_param_count = int(t.value)
        self._loss = torch.zeros((1,), dtype=torch.float32, device=device.torch_device)

    @classmethod
    def for_scene(cls, device: Device, scene, config: RadianceCacheConfig | None = None, margin: float = 0.01):
        """A cache over the scene's world box, grown by ``margin`` of its
        extent on every side."""
        lo, hi = scene.bounds()
        ext = np.maximum(hi - lo, 1e-3)
        return cls(device, lo - margin * ext, hi + margin * ext, config)

    def close(self) -> None:
        if getattr(self, "handle", None):
            _native.lib().nan_nrc_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters ---------------------------------------------------------
    def get_params(self) -> tuple[np.ndarray, np.ndarray]:
        """FP32 master parameters: weights float32[16384] (per layer [out][in])
        and table float32[levels, entries, 2]."""
        w = np.empty(self.weight_count, dtype=np.float32)
        t = np.empty(self.table_param_count, dtype=np.float32)
        _native.check(_native.lib().nan_nrc_get_params(self.handle, w.ctypes.data, t.ctypes.data),
                      "nan_nrc_get_params")
        return w, t.reshape(self.config.levels, -1, 2)

    def set_params(self, weights: np.ndarray | None = None, table: np.ndarray | None = None) -> None:
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32).reshape(-1)
        t = None if table is None else np.ascontiguousarray(table, dtype=np.float32).reshape(-1)
        if w is not None and w.size != self.weight_count:
            raise ValueError(f"expected {self.weight_count} weights")
        if t is not None and t.size != self.table_param_count:
            raise ValueError(f"expected {self.table_param_count} table parameters")
        _native.check(_native.lib().nan_nrc_set_params(self.handle, None if w is None else w.ctypes.data,
                                                       None if t is None else t.ctypes.data),
                      "nan_nrc_set_params")

    @staticmethod
    def split_weights(weights: np.ndarray) -> list[np.ndarray]:
        """Per-layer [out, in] views of a flat weight vector."""
        out, off = [], 0
        for o, i in LAYER_SHAPES:
            out.append(weights[off:off + o * i].reshape(o, i))
            off += o * i
        return out

    # -- queries ------------------------------------------------------------
    def _inputs(self, pos, nrm, alb):
        n = pos.shape[0]
        dev = self.device.torch_device
        for t, name in ((pos, "pos"), (nrm, "nrm"), (alb, "alb")):
            _check_rows(t, n, 3, name, dev)
        return n

    def _stream(self, encoder: CommandEncoder | None) -> int | None:
        return (encoder or self.device.create_command_encoder()).stream_handle

    def infer(self, pos: torch.Tensor, nrm: torch.Tensor, alb: torch.Tensor,
              encoder: CommandEncoder | None = None) -> torch.Tensor:
        """Predicted RGB radiance (clamped at 0) float32[n, 3]."""
        n = self._inputs(pos, nrm, alb)
        out = torch.empty((n, 3), dtype=torch.float32, device=self.device.torch_device)
        _native.check(_native.lib().nan_nrc_infer(self.handle, _ptr(pos), _ptr(nrm), _ptr(alb), _ptr(out), n,
                                                  self._stream(encoder)), "nan_nrc_infer")
        return out

    def forward_raw(self, pos, nrm, alb, encoder: CommandEncoder | None = None) -> torch.Tensor:
        """The first four raw network outputs float32[n, 4]."""
        n = self._inputs(pos, nrm, alb)
        out = torch.empty((n, 4), dtype=torch.float32, device=self.device.torch_device)
        _native.check(_native.lib().nan_nrc_forward_raw(self.handle, _ptr(pos), _ptr(nrm), _ptr(alb), _ptr(out), n,
                                                        self._stream(encoder)), "nan_nrc_forward_raw")
        return out

    def encode(self, pos, nrm, alb, encoder: CommandEncoder | None = None) -> torch.Tensor:
        """Network input float32[n, 48] (hash features, normal, albedo, zeros)."""
        n = self._inputs(pos, nrm, alb)
        out = torch.empty((n, INPUT_WIDTH), dtype=torch.float32, device=self.device.torch_device)
        _native.check(_native.lib().nan_nrc_encode(self.handle, _ptr(pos), _ptr(nrm), _ptr(alb), _ptr(out), n,
                                                   self._stream(encoder)), "nan_nrc_encode")
        return out

    def time_forward(self, pos, nrm, alb, repeats: int = 10, encoder: CommandEncoder | None = None) -> float:
        """Milliseconds per launch of the fused MLP forward kernel alone on
        these (already encoded) queries, from HIP events on the stream."""
        n = self._inputs(pos, nrm, alb)
        ms = ctypes.c_float()
        _native.check(_native.lib().nan_nrc_time_forward(self.handle, _ptr(pos), _ptr(nrm), _ptr(alb), n,
                                                         int(repeats), ctypes.byref(ms), self._stream(encoder)),
                      "nan_nrc_time_forward")
        return float(ms.value)

    def backward_raw(self, pos, nrm, alb, target, encoder: CommandEncoder | None = None):
        """Gradients of the mean loss: (weights float32[16384], inputs
        float32[n, 48]); parameters are not updated."""
        n = self._inputs(pos, nrm, alb)
        _check_rows(target, n, 3, "target", self.device.torch_device)
        dev = self.device.torch_device
        wg = torch.empty((self.weight_count,), dtype=torch.float32, device=dev)
        ig = torch.empty((n, INPUT_WIDTH), dtype=torch.float32, device=dev)
        _native.check(_native.lib().nan_nrc_backward_raw(self.handle, _ptr(pos), _ptr(nrm), _ptr(alb), _ptr(target),
                                                         n, _ptr(wg), _ptr(ig), self._stream(encoder)),
                      "nan_nrc_backward_raw")
        return wg, ig

    def table_grad_raw(self, pos, nrm, alb, target, encoder: CommandEncoder | None = None) -> torch.Tensor:
        """The hash-table gradient of the mean loss as the Adam step receives
        it (the packed FP16 atomic sums, rescaled), float32[levels, entries,
        2]; parameters are not updated.  Test hook."""
        n = self._inputs(pos, nrm, alb)
        _check_rows(target, n, 3, "target", self.device.torch_device)
        out = torch.empty((self.table_param_count,), dtype=torch.float32, device=self.device.torch_device)
        _native.check(_native.lib().nan_nrc_table_grad(self.handle, _ptr(pos), _ptr(nrm), _ptr(alb), _ptr(target),
                                                       n, _ptr(out), self._stream(encoder)), "nan_nrc_table_grad")
        return out.view(self.config.levels, -1, 2)

    def set_fused_training(self, fused: bool) -> None:
        """Training kernels: the three-launch fused step (default; minibatches
        of up to 32,768 records, dense and hashed levels) or the seven-launch
        step with packed FP16 table-gradient atomics (larger minibatches
        always take that one)."""
        _native.check(_native.lib().nan_nrc_set_fused_training(self.handle, 1 if fused else 0),
                      "nan_nrc_set_fused_training")

    def last_step_stats(self) -> tuple[float, float]:
        """(mean loss, valid records) of the last Adam step; waits for the device."""
        out = (ctypes.c_float * 2)()
        _native.check(_native.lib().nan_nrc_stats(self.handle, out), "nan_nrc_stats")
        return float(out[0]), float(out[1])

    def train_step(self, pos, nrm, alb, target, encoder: CommandEncoder | None = None) -> torch.Tensor:
        """Adam step(s) on the records (``config.minibatches`` of them);
        returns the last step's mean loss as a device scalar."""
        n = self._inputs(pos, nrm, alb)
        _check_rows(target, n, 3, "target", self.device.torch_device)
        self._loss.zero_()
        _native.check(_native.lib().nan_nrc_train(self.handle, _ptr(pos), _ptr(nrm), _ptr(alb), _ptr(target), n,
                                                  _ptr(self._loss), self._stream(encoder)), "nan_nrc_train")
        return self._loss.clone()
