"""Many independent scenes traced together: one set of wavefront launches
(and, optionally, one captured hipGraph) per frame for all of them.

Each scene keeps its own BVH (rebased into shared arrays), Mesh Colors
layout, cache, lighting, RNG streams and path length, so every slot equals a
one-at-a-time run of that scene.  The reference traces one scene and
dispatches per instance (texture_space_path_tracer.py:42-59); batching has no
reference counterpart.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native
from .device import CommandEncoder, Device
from .mesh_colors import MeshColorsLayout
from .scene import Scene
from .scene_node import SceneNode
from .texture_space_path_tracer import TextureSpacePathTracer


class SceneBatch:
    """Scenes, layouts and caches of a batch.

    ``layout_kwargs`` is one dict of ``MeshColorsLayout.build`` arguments per
    scene; ``max_bounces`` one path length per scene (1..8 in the batched
    workload)."""

    def __init__(self, device: Device, nodes: list[SceneNode], layout_kwargs: list[dict],
                 max_bounces: list[int], samples_per_texel: int = 1, rng: str = "reference"):
        if not (len(nodes) == len(layout_kwargs) == len(max_bounces)) or not nodes:
            raise ValueError("one layout and one path length per scene")
        self.device = device
        self.nodes = nodes
        self.max_bounces = [max(1, int(b)) for b in max_bounces]
        self.samples_per_texel = max(1, int(samples_per_texel))
        if rng not in _native.RNG_MODES:
            raise ValueError(f"unknown rng {rng!r}")
        self.rng = rng
        self.layouts = [MeshColorsLayout.build(n, **kw) for n, kw in zip(nodes, layout_kwargs)]
        self.scenes = [Scene(device, n) for n in nodes]
        self.tracers = [TextureSpacePathTracer(device, s, l, samples_per_texel=self.samples_per_texel,
                                               max_bounces=b)
                        for s, l, b in zip(self.scenes, self.layouts, self.max_bounces)]
        dev = device.torch_device
        self.payloads = [torch.zeros((l.total_payload_count, 4), dtype=torch.float32, device=dev)
                         for l in self.layouts]
        n = len(nodes)
        scene_h = (ctypes.c_void_p * n)(*[s.handle.value for s in self.scenes])
        tracer_h = (ctypes.c_void_p * n)(*[t.handle.value for t in self.tracers])
        handle = ctypes.c_void_p()
        err = _native.err_buffer()
        rc = _native.lib().nan_batch_create(scene_h, tracer_h, n, ctypes.byref(handle), err, len(err))
        if rc != 0:
            raise _native.NanError(err.value.decode())
        self.handle = handle
        ptrs = (ctypes.c_void_p * n)(*[p.data_ptr() for p in self.payloads])
        _native.check(_native.lib().nan_batch_set_payloads(self.handle, ptrs), "nan_batch_set_payloads")
        b = np.asarray(self.max_bounces, dtype=np.uint32)
        _native.check(_native.lib().nan_batch_set_bounces(self.handle, b.ctypes.data_as(_native.u32p)),
                      "nan_batch_set_bounces")
        self.tuning = (0, 16, 16)
        self.graph = None

    def close(self) -> None:
        if getattr(self, "handle", None):
            _native.lib().nan_tracer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def total_slots(self) -> int:
        return sum(l.total_payload_count for l in self.layouts)

    @property
    def triangle_counts(self) -> list[int]:
        return [n.triangle_count for n in self.nodes]

    def tune(self, blocks_per_cu: int = 0, lds_stack: int = 16, refill_steps: int = 16) -> None:
        _native.check(_native.lib().nan_tracer_tune(self.handle, int(blocks_per_cu), int(lds_stack),
                                                    int(refill_steps)), "nan_tracer_tune")
        self.tuning = (int(blocks_per_cu), int(lds_stack), int(refill_steps))
        self.graph = None

    def execute(self, command_encoder: CommandEncoder, iteration: int, *, reset: bool = False,
                counters: torch.Tensor | None = None) -> None:
        """One iteration of every scene (all slots of all caches)."""
        stream = command_encoder.stream_handle
        mb = max(self.max_bounces)
        if counters is None:
            rc = _native.lib().nan_tst_execute(self.handle, None, int(iteration), self.samples_per_texel, mb,
                                               1 if reset else 0, _native.RNG_MODES[self.rng], stream)
        else:
            rc = _native.lib().nan_tst_execute_counted(self.handle, None, int(iteration), self.samples_per_texel, mb,
                                                       1 if reset else 0, _native.RNG_MODES[self.rng],
                                                       counters.data_ptr(), stream)
        _native.check(rc, "nan_tst_execute (batch)")

    def set_frame_index(self, command_encoder: CommandEncoder, frame_index: int) -> None:
        """Frame index of the blend alpha test, shared by every scene of the
        batch (a one-at-a-time run of a scene with the same frame index gives
        the same slots)."""
        _native.check(_native.lib().nan_batch_set_frame_index(self.handle, int(frame_index) & 0xFFFFFFFF,
                                                              command_encoder.stream_handle),
                      "nan_batch_set_frame_index")

    def profile(self, enable: bool) -> None:
        _native.check(_native.lib().nan_tracer_profile(self.handle, 1 if enable else 0), "nan_tracer_profile")

    def stage_times(self) -> dict:
        ms = (ctypes.c_double * len(_native.STAGES))()
        n = (ctypes.c_uint32 * len(_native.STAGES))()
        _native.check(_native.lib().nan_tracer_stage_times(self.handle, ms, n), "nan_tracer_stage_times")
        return {name: (ms[k], n[k]) for k, name in enumerate(_native.STAGES) if n[k]}

    def set_device_frame(self, iteration: int, reset: bool = False) -> None:
        """Iteration/reset come from a device counter advanced by each call."""
        _native.check(_native.lib().nan_tracer_device_frame(self.handle, 1, int(iteration), 1 if reset else 0),
                      "nan_tracer_device_frame")

    def clear_device_frame(self) -> None:
        _native.check(_native.lib().nan_tracer_device_frame(self.handle, 0, 0, 0), "nan_tracer_device_frame")

    def capture(self, command_encoder: CommandEncoder, next_iteration: int,
                reset: bool = False) -> "torch.cuda.CUDAGraph":
        """Captures one batched frame into a hipGraph.  Replays run iterations
        next_iteration, next_iteration + 1, ..., the first of them resetting
        the caches when ``reset``.  The working buffers exist from the
        batch's creation (and from the last ``tune``), so this may be the
        batch's first frame."""
        self.set_device_frame(next_iteration, reset)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.execute(command_encoder, 0)
        self.graph = graph
        return graph


def batch_workload(seed: int = 0, count: int = 64, *, tri_range=(1_000, 200_000),
                   budget_range=(16_384, 2_097_152), bounce_range=(1, 8), roulette: bool = True):
    """Seeded many-scene workload: triangle counts and texel budgets
    log-uniform in the given ranges.  With ``roulette`` every scene runs
    Russian roulette under a cap of bounce_range[1] bounces, so its paths end
    after bounce_range[0]..bounce_range[1] bounces; without it each scene's
    fixed path length is drawn uniformly from bounce_range (the draw is made
    either way, so the scenes themselves do not depend on the choice).
    Per scene the layout uses texels_per_unit 8 (halved until the budget can
    hold the scene), max_resolution 16 and the largest min_resolution in
    {4, 2, 1} whose unhalved demand fits the budget.  A drawn budget below
    the scene's minimum (3 slots per triangle side, i.e. min_resolution 1) is
    raised to that minimum.  Returns (nodes,
    layout_kwargs, bounces)."""
    from .scenes import batch_scene

    rng = np.random.default_rng(seed)
    nodes, kws, bounces = [], [], []
    for k in range(count):
        tris = int(np.exp(rng.uniform(np.log(tri_range[0]), np.log(tri_range[1]))))
        budget = int(np.exp(rng.uniform(np.log(budget_range[0]), np.log(budget_range[1]))))
        node = batch_scene(int(rng.integers(0, 2**31)), tris)
        from .mesh_colors import _resolutions, instance_max_edges

        edges = instance_max_edges(node)
        mult = np.concatenate([np.full(len(e), 2 if node.materials[m].double_sided else 1, dtype=np.int64)
                               for e, (_, m, _) in zip(edges, node.instances)])
        edges = np.concatenate(edges)

        def demand(tpu: float, min_res: int) -> int:
            res = _resolutions(edges, tpu, min_res, 16)
            return int((((res + 1) * (res + 2) // 2) * mult).sum())

        budget = max(budget, int(3 * mult.sum()))  # the scene's minimum: 3 slots per triangle side
        kw = None
        tpu = 8.0
        while kw is None:
            for min_res in (4, 2, 1):
                if demand(tpu, min_res) <= budget:
                    kw = dict(texels_per_unit=tpu, min_resolution=min_res, max_resolution=16,
                              max_total_texels=budget)
                    break
            tpu *= 0.5
        drawn = int(rng.integers(bounce_range[0], bounce_range[1] + 1))
        node.russian_roulette = roulette
        nodes.append(node)
        kws.append(kw)
        bounces.append(int(bounce_range[1]) if roulette else drawn)
    return nodes, kws, bounces
