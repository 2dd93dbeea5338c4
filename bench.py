"""Benchmark: the texture-space path tracer on one MI355X.

The metric is effective converged texel-path samples/s at equal RMSE
against a 4096-frame plain bake, learned versus plain.  The default run
(``python bench.py``) measures it on the configuration it is defined on: the
Cornell box (36 triangles, texels_per_unit 16, res 4-64, 1 spp, 3 bounces)
with the learned radiance cache in its specified form (65,536 training
suffixes per frame, paths ending in the cache after one bounce).  ``value``
is the plain samples that reach the learned RMSE after 256 frames over the
learned wall time, the learned frames priced by exactly ``--steps`` timed
frames (after ``--warmup`` more); ``value_vs_plain`` is its ratio to the
plain tracer's own samples/s.

Beside it, in the same JSON line:

* ``plain_at_scale``: the plain tracer on the seeded synthetic 1.21M-triangle
  scene (512x512-quad value-noise heightfield, 64x64 extruded boxes, 500
  level-3 icospheres; sun at 35 deg + sky 0.3; texels_per_unit 8, res 2-16,
  budget 16,777,216), ``--steps`` iterations timed with HIP events on the
  launch stream around the production path (the double-sided texels' paths
  on a second stream beside the rest), then a profiled replay with HIP
  events around every tracer launch (one-stream order, so every launch is
  timed alone) and a counted replay for the work counts.
* ``roofline``: the dominant kernel of that leg, the closest-hit traversal
  ``wf_extend``: algorithmic bytes (64 B per counted node visit, 48 B per
  triangle test, the queued ray and hit record) over its launch time in the
  profiled replay, against 8 TB/s; with HBM traffic, L2 and issue counters
  from the committed rocprofv3 summary, and ``bound`` naming the resource
  closest to saturation.
* ``learned_vs_plain_at_scale``: both learned forms on the 1.21M-triangle
  scene against a 4096-frame bake (the texel-query form is a biased variant
  outside the specification).
* ``batch``: 64 seeded scenes (3.06M triangles, 21.4M slots, Russian
  roulette paths of 1-8 bounces) in one set of launches per frame, every
  frame a replayed hipGraph, the launch shape autotuned first.
* ``combined``: the 4.0M-triangle scene as a batch of one under hipGraph
  replay (latency percentiles, Mrays/s), a 3840x2160 resolve, the learned
  forms against a 4096-frame bake.
* ``cpu_baseline``: the oracle tracer on a strided sample of the
  1.21M-triangle scene on the host's cores, and the reference's own CPU
  ray path (tinybvh) beside it.

``--workload cornell|batch|learned|combined`` runs one leg alone.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# bumped whenever the tracer kernels change, so stale rocprof summaries are ignored
KERNEL_VERSION = 6
METRIC = "texel-path samples/s (Mrays/s) at equal RMSE vs 4096-spp bake; learned vs plain"

# algorithmic bytes (see the module docstring)
BYTES_PER_NODE = 64
BYTES_PER_TRI = 48
BYTES_PER_SLOT_UPDATE = 32         # 16-byte slot read + write
BYTES_PER_TEXEL_SETUP = 4 + 4 + 16 + 8 + 4 + 8   # texel->face, face->instance, instance, side, prefix, face
# per shaded hit: 80-byte triangle shading record + transform rows 48 + inverse
# transpose 64 + material 32
BYTES_PER_SHADED_HIT = 80 + 48 + 64 + 32
BYTES_PER_RAY = 32                 # queued origin + direction (float4 x 2)
BYTES_PER_HIT = 16                 # t, b1, b2, triangle id
BYTES_PER_SHADOW_RAY = 48          # origin, direction, contribution (float4 x 3)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str):
    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.scenes import cornell_box, synthetic_scene

    if name == "cornell":
        node = cornell_box()
        kw = dict(texels_per_unit=16.0, min_resolution=4, max_resolution=64, max_total_texels=16_777_216)
        desc = ("plain texel path tracer, Cornell box (36 tris), texels_per_unit 16, res 4-64, "
                "1 spp, 3 bounces, FP32")
    else:
        node = synthetic_scene(0)
        kw = dict(texels_per_unit=8.0, min_resolution=2, max_resolution=16, max_total_texels=16_777_216)
        desc = ("plain texel path tracer, synthetic 1.21M-tri scene seed 0 (heightfield + 4096 boxes + "
                "500 icospheres, sun 35 deg + sky 0.3), texels_per_unit 8, res 2-16, budget 16,777,216, "
                "1 spp, 3 bounces, FP32")
    layout = MeshColorsLayout.build(node, **kw)
    return node, layout, kw, desc


def cpu_baseline(node, layout, threads: int, target_seconds: float = 12.0) -> dict:
    """Oracle tracer (C, pthreads) on a strided sample of front texels."""
    import numpy as np

    import oracle
    from nan_amd.scene_arrays import flatten_scene

    t0 = time.perf_counter()
    orc = oracle.OracleScene(flatten_scene(node))
    build_s = time.perf_counter() - t0
    payload = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    total = layout.total_surface_texels
    stride = max(1, total // 4096)
    work = np.arange(0, total, stride, dtype=np.uint64)
    t0 = time.perf_counter()
    cnt = orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, payload, 0, 1, 3, True,
                           work=work, threads=threads)
    dt = time.perf_counter() - t0
    # scale the sample so the timed run is ~target_seconds
    per_texel = dt / max(1, len(work))
    n = int(min(total, max(len(work), target_seconds / max(per_texel, 1e-9))))
    stride = max(1, total // n)
    work = np.arange(0, total, stride, dtype=np.uint64)
    t0 = time.perf_counter()
    cnt = orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, payload, 0, 1, 3, True,
                           work=work, threads=threads)
    dt = time.perf_counter() - t0
    samples = int(cnt[3])
    rays = int(cnt[0] + cnt[1])
    return {
        "value": samples / dt,
        "unit": "texel-path samples/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle C tracer, iteration 0, every {stride}th front texel ({len(work)} texels, "
                   f"{samples} slot updates, {rays} rays) in {dt:.1f} s; oracle BVH build {build_s:.1f} s"),
        "mrays_per_s": rays / dt / 1e6,
    }


def tinybvh_baseline(node, threads: int, rays: int = 4_000_000) -> dict | None:
    """The reference's own CPU ray path, timed: the tinybvh it vendors
    (vertex_baker/cpp/third_party/tinybvh, v1.8.0), binned-SAH build and
    IsOccluded per ray on one worker per thread (tinybvh_visibility.cpp:
    145-230), through oracle/_ref/libtinybvh_ref.so, on sun shadow rays from
    uniformly drawn points of the scene's triangles.  None when the harness is
    not built."""
    import ctypes

    import numpy as np

    import oracle
    from nan_amd.scene_arrays import flatten_scene

    lib_path = ROOT / "oracle" / "_ref" / "libtinybvh_ref.so"
    if not lib_path.exists():
        return None
    lib = ctypes.CDLL(str(lib_path))
    lib.ref_bvh_create.restype = ctypes.c_void_p
    lib.ref_bvh_create.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    lib.ref_bvh_build_ms.restype = ctypes.c_double
    lib.ref_bvh_build_ms.argtypes = [ctypes.c_void_p]
    lib.ref_bvh_destroy.argtypes = [ctypes.c_void_p]
    lib.ref_bvh_occluded.restype = ctypes.c_double
    lib.ref_bvh_occluded.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int]
    arrays = flatten_scene(node)
    world = np.ascontiguousarray(oracle.OracleScene(arrays, use_bvh=False).world_triangles(), dtype=np.float32)
    tris = world.reshape(-1, 3, 3).astype(np.float64)
    rng = np.random.default_rng(0)
    pick = rng.integers(0, len(tris), size=rays)
    a, b = rng.random(rays), rng.random(rays)
    flip = a + b > 1.0
    a[flip], b[flip] = 1.0 - a[flip], 1.0 - b[flip]
    t = tris[pick]
    p = t[:, 0] + a[:, None] * (t[:, 1] - t[:, 0]) + b[:, None] * (t[:, 2] - t[:, 0])
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    sun = np.asarray(arrays["lighting"][0:3], dtype=np.float64)
    n *= np.where((n @ sun) < 0.0, -1.0, 1.0)[:, None]
    r = np.zeros((rays, 8), dtype=np.float32)
    r[:, 0:3] = p + 1e-3 * n
    r[:, 4:7] = sun
    r[:, 7] = 1000.0
    occ = np.zeros(rays, dtype=np.uint32)
    h = lib.ref_bvh_create(world.ctypes.data, len(tris))
    try:
        build_ms = lib.ref_bvh_build_ms(h)
        trace_ms = lib.ref_bvh_occluded(h, r.ctypes.data, rays, occ.ctypes.data, threads)
    finally:
        lib.ref_bvh_destroy(h)
    return {"value": rays / (trace_ms * 1e-3) / 1e6, "unit": "Mrays/s (occlusion)", "cores": threads,
            "kind": "reference", "build_ms": build_ms, "trace_ms": trace_ms,
            "sample": (f"{rays} sun shadow rays from uniformly drawn triangle points (offset 1e-3 along the "
                       f"normal), {float(occ.mean()):.3f} occluded; tinybvh v1.8.0 binned-SAH BVH over "
                       f"{len(tris)} triangles")}


MFMA_F16_PEAK_TFLOPS = 2500.0     # dense FP16 MFMA peak of MI355X
# The network's products run on v_mfma_f32_16x16x16_f16 (512 FLOP per SIMD
# cycle, about half the dense peak): the doubled-K form gave wrong products
# where the compiler rewrote its operand registers shortly after issue.
MFMA_INSTRUCTION = "v_mfma_f32_16x16x16_f16"
MFMA_K16_PEAK_TFLOPS = 512 * 1024 * 2.4e9 / 1e12  # 1,024 SIMDs at 2.4 GHz
MLP_FLOPS_PER_QUERY = 2 * (48 * 64 + 3 * 64 * 64 + 64 * 16)   # executed (padded) forward FLOPs: 32,768


def _plain_frames_for(curve: dict, r: float) -> float:
    """Plain frames that reach RMSE r: log-log interpolation of the measured
    plain curve, extended beyond its ends with the 1/sqrt(frames) law."""
    import math

    import numpy as np

    ks = np.array(sorted(curve), dtype=np.float64)
    rs = np.array([curve[k] for k in sorted(curve)], dtype=np.float64)
    if r >= rs[0]:
        return float(ks[0] * (rs[0] / r) ** 2)
    if r <= rs[-1]:
        return float(ks[-1] * (rs[-1] / r) ** 2)
    return float(np.exp(np.interp(-math.log(r), -np.log(rs), np.log(ks))))


def cache_config(scene, layout, minibatches: int):
    """The radiance cache's grid matched to the layout's texel density
    (RadianceCacheConfig.for_texel_density over the scene's longest extent)."""
    import numpy as np

    from nan_amd.radiance_cache import RadianceCacheConfig

    lo, hi = scene.bounds()
    return RadianceCacheConfig.for_texel_density(float(np.max(hi - lo)), layout.texels_per_unit,
                                                 minibatches=minibatches)


# Adam steps per frame of the Cornell-box learned form, trained beside the
# next frame's tracing.  Against the plain tracer at equal RMSE after 256
# frames (two cache seeds each, one run varies by about +-0.1): 4 steps of
# 16,384 records (0.22 ms per frame) 0.46-0.68x; 3 steps 0.44-0.59x; 2 steps
# of 32,768 (0.18-0.20 ms) 0.29-0.61x.  More, smaller steps converge in fewer
# frames; the split of the 65,536 records barely changes the frame's cost.
LEARNED_MINIBATCHES = 4


def learned_measurement(device, frames=(16, 64, 256), bake_frames: int = 4096, plain_curve_frames: int = 2048,
                        train_samples: int = 65536, mlp_queries: int = 1 << 22, steps: int = 10,
                        warmup: int = 2) -> dict:
    """Learned radiance cache against the plain tracer on the Cornell box
    (texels_per_unit 16, res 4-64, 1 spp, 3 bounces), the specified form
    (paths end in the cache after one bounce): RMSE of both against a plain
    bake after 16, 64 and 256 frames, the plain frames that reach each
    learned RMSE and the effective converged texel-path samples/s (those
    plain samples over the learned wall time).  After the RMSE run,
    `warmup` more learned frames, then exactly `steps` timed ones: their
    mean frame time prices the learned frames in `value`.  Also the fused
    MLP forward kernel's MFMA rate on a large query batch."""
    import numpy as np
    import torch

    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scene import Scene
    from nan_amd.scenes import cornell_box
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    node = cornell_box()
    layout = MeshColorsLayout.build(node, texels_per_unit=16.0, min_resolution=4, max_resolution=64)
    scene = Scene(device, node)
    tracer = TextureSpacePathTracer(device, scene, layout, samples_per_texel=1, max_bounces=3, mode="megakernel")
    enc = device.create_command_encoder()
    stream = torch.cuda.current_stream(device.torch_device)
    slots = layout.total_payload_count
    dev = device.torch_device
    bake = torch.zeros((slots, 4), dtype=torch.float32, device=dev)
    for it in range(bake_frames):  # its own RNG streams: iterations 1,000,000 + k
        tracer.execute(enc, bake, 1_000_000 + it, reset=(it == 0))
    b = bake[:, :3].double()

    def rmse(p):
        return float(torch.sqrt(torch.mean((p[:, :3].double() - b) ** 2)))

    plain = torch.zeros_like(bake)
    curve = {}
    for it in range(plain_curve_frames):
        tracer.execute(enc, plain, it, reset=(it == 0))
        if (it + 1) & it == 0:
            curve[it + 1] = rmse(plain)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for it in range(512):
        tracer.execute(enc, plain, it, reset=(it == 0))
    ev1.record(stream)
    torch.cuda.synchronize()
    plain_ms = ev0.elapsed_time(ev1) / 512

    cfg = cache_config(scene, layout, LEARNED_MINIBATCHES)
    cache = RadianceCache.for_scene(device, scene, cfg)
    learned = torch.zeros_like(bake)
    out_frames = {}
    elapsed_ms = 0.0
    done = 0
    tracer.profile(True)
    for target in frames:
        ev0.record(stream)
        for it in range(done, target):
            tracer.execute_learned(enc, learned, it, cache, reset=(it == 0), train_samples=train_samples,
                                   overlap=True)
        ev1.record(stream)
        torch.cuda.synchronize()
        elapsed_ms += ev0.elapsed_time(ev1)
        done = target
        r = rmse(learned)
        m = _plain_frames_for(curve, r)
        learned_ms = elapsed_ms / target
        out_frames[str(target)] = {
            "rmse": r,
            "plain_frames_for_equal_rmse": m,
            "learned_ms_per_frame": learned_ms,
            "effective_samples_per_s": m * slots / (elapsed_ms * 1e-3),
            "speedup_vs_plain": (m * plain_ms) / elapsed_ms,
        }
    stages = tracer.stage_times()
    tracer.profile(False)
    # the timed steps: learned frames continuing the run above
    it = frames[-1]
    for _ in range(warmup):
        tracer.execute_learned(enc, learned, it, cache, train_samples=train_samples, overlap=True)
        it += 1
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(steps):
        tracer.execute_learned(enc, learned, it, cache, train_samples=train_samples, overlap=True)
        it += 1
    ev1.record(stream)
    torch.cuda.synchronize()
    timed_ms = ev0.elapsed_time(ev1) / steps
    # fused MLP forward alone on a large batch of queries inside the scene box
    rng = np.random.default_rng(5)
    lo, hi = scene.bounds()
    q = mlp_queries
    pos = torch.from_numpy(rng.uniform(lo, hi, size=(q, 3)).astype(np.float32)).to(dev)
    nrm = torch.nn.functional.normalize(torch.randn((q, 3), device=dev), dim=1).contiguous()
    alb = torch.rand((q, 3), device=dev)
    mlp_ms = cache.time_forward(pos, nrm, alb, repeats=20, encoder=enc)
    tflops = MLP_FLOPS_PER_QUERY * q / (mlp_ms * 1e-3) / 1e12
    last = out_frames[str(frames[-1])]
    return {
        "workload": ("Cornell box (36 tris), texels_per_unit 16, res 4-64, 1 spp, 3 bounces; learned cache: "
                     f"{cfg.levels}-level hash grid (16 to {cfg.max_resolution:g} cells, matched to the texel "
                     "density) x 2 FP16 features, 2^19 entries/level, normal + albedo, 4x64 ReLU MLP "
                     f"FP16 MFMA / FP32 accumulate, Adam lr 1e-2, relative L2, {train_samples} training suffixes "
                     f"per frame in {LEARNED_MINIBATCHES} Adam steps, trained on a second stream beside the next "
                     "frame's tracing (each frame's queries see the cache trained through the previous frame), "
                     "paths end in the cache after 1 bounce, emitter light sampling with MIS, ramped mean"),
        "cache_levels": cfg.levels,
        "cache_max_resolution": cfg.max_resolution,
        "bake": f"{bake_frames}-frame plain bake (independent RNG streams)",
        "payload_slots": slots,
        "plain_ms_per_frame": plain_ms,
        "plain_samples_per_s": slots / (plain_ms * 1e-3),
        "plain_rmse": {str(k): v for k, v in curve.items()},
        "learned": out_frames,
        "form": "one bounce, the specified estimator: E = sun + NEE(emitters, MIS) + (az/pdf)(Le(y1) + "
                "max(0, cache(y1)))",
        "timed_ms_per_frame": timed_ms,
        "value_frames": frames[-1],
        "bake_frames": bake_frames,
        "value": last["plain_frames_for_equal_rmse"] * slots / (frames[-1] * timed_ms * 1e-3),
        "speedup_vs_plain": last["plain_frames_for_equal_rmse"] * plain_ms / (frames[-1] * timed_ms),
        "learned_stage_ms_per_frame": {k: v[0] / frames[-1] for k, v in stages.items()},
        "mlp_roofline": {
            "bound": "mfma", "kernel": "nrc_mlp_fwd<false>", "queries": q, "avg_launch_ms": mlp_ms,
            "flops_per_query": MLP_FLOPS_PER_QUERY, "achieved": tflops, "peak": MFMA_F16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": tflops / MFMA_F16_PEAK_TFLOPS,
            "instruction": MFMA_INSTRUCTION, "instruction_peak": MFMA_K16_PEAK_TFLOPS,
            "instruction_frac": tflops / MFMA_K16_PEAK_TFLOPS,
        },
    }


LEARNED_FORMS = {
    # name: (texel_query, training records per frame, Adam minibatches)
    "one_bounce": (False, 65536, 8),
    "texel_query": (True, 131072, 16),
}
LEARNED_FORM_NOTES = {
    "one_bounce": "the specified estimator: paths end in the cache after one bounce (unbiased for a converged "
                  "cache)",
    "texel_query": "a biased variant outside the specification: paths end in the cache at the texel itself (no "
                   "bounce ray), so the fit's bias stays in every frame",
}
LEARNED_FRAMES = (1, 2, 4, 8, 16, 64)


def learned_forms_vs_plain(device, scene, tracer, rmse, curve: dict, plain_ms: float, frames=LEARNED_FRAMES,
                           graph: bool = False) -> dict:
    """Each form of LEARNED_FORMS from a fresh cache: its RMSE (``rmse``
    against the plain bake) after each frame count in ``frames``, the plain
    frames that reach it (``curve``: plain RMSE per frame count) and the
    equal-RMSE speedup (plain frames x ``plain_ms`` over the learned time),
    plus the learned frame's stage times.  With ``graph`` every learned frame
    is a replay of one captured iteration in device-frame mode (first frame
    included), and the per-frame latency percentiles are reported."""
    import numpy as np
    import torch

    from nan_amd.radiance_cache import RadianceCache

    enc = device.create_command_encoder()
    stream = torch.cuda.current_stream(device.torch_device)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    slots = tracer.layout.total_payload_count
    forms = {}
    for name, (texel_query, n_train, mb) in LEARNED_FORMS.items():
        cfg = cache_config(scene, tracer.layout, mb)
        cache = RadianceCache.for_scene(device, scene, cfg)
        learned = torch.zeros((slots, 4), dtype=torch.float32, device=device.torch_device)
        res = {}
        elapsed = 0.0
        lat = []
        g = None
        if graph:
            g = tracer.capture_learned(enc, learned, cache, 0, reset=True, train_samples=n_train,
                                       texel_query=texel_query)
        for it in range(max(frames)):
            ev0.record(stream)
            if g is not None:
                g.replay()
            else:
                tracer.execute_learned(enc, learned, it, cache, reset=(it == 0), train_samples=n_train,
                                       texel_query=texel_query)
            ev1.record(stream)
            torch.cuda.synchronize()
            lat.append(ev0.elapsed_time(ev1))
            elapsed += lat[-1]
            if it + 1 in frames:
                r = rmse(learned)
                m = _plain_frames_for(curve, r)
                res[str(it + 1)] = {"rmse": r, "plain_frames_for_equal_rmse": m,
                                    "learned_ms_per_frame": elapsed / (it + 1),
                                    "speedup_vs_plain": m * plain_ms / elapsed}
        if g is not None:
            tracer.clear_device_frame()
            del g
        tracer.profile(True)
        for it in range(2):
            tracer.execute_learned(enc, learned, max(frames) + it, cache, train_samples=n_train,
                                   texel_query=texel_query)
        stages = {k: v[0] / 2 for k, v in tracer.stage_times().items()}
        tracer.profile(False)
        forms[name] = {"form": LEARNED_FORM_NOTES[name], "train_samples": n_train, "minibatches": mb,
                       "cache_levels": cfg.levels, "frames": res, "stage_ms_per_frame": stages}
        if graph:
            forms[name]["graph_replays"] = True
            forms[name]["latency_ms"] = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))}
        log(f"learned {name}: " + ", ".join(f"{k}: x{v['speedup_vs_plain']:.2f}" for k, v in res.items()))
        del cache, learned
    return forms


def learned_at_scale(device, scene, layout, tracer, bake_frames: int = 4096, frames=LEARNED_FRAMES) -> dict:
    """Both learned forms against the plain tracer on the workload's own
    scene at equal RMSE against a plain bake (``learned_forms_vs_plain``).
    ``one_bounce`` is the configured form (65,536 suffixes per frame, paths
    end in the cache after one bounce); ``texel_query`` queries the cache at
    the texel itself (no bounce ray; 131,072 training paths per frame in 16
    Adam steps)."""
    import torch

    enc = device.create_command_encoder()
    stream = torch.cuda.current_stream(device.torch_device)
    slots = layout.total_payload_count
    dev = device.torch_device
    bake = torch.zeros((slots, 4), dtype=torch.float32, device=dev)
    for it in range(bake_frames):  # its own RNG streams: iterations 1,000,000 + k
        tracer.execute(enc, bake, 1_000_000 + it, reset=(it == 0))
        if it % 512 == 511:
            torch.cuda.synchronize()
            log(f"scale bake {it + 1}/{bake_frames}")
    b = bake[:, :3].double()

    def rmse(p):
        return float(torch.sqrt(torch.mean((p[:, :3].double() - b) ** 2)))

    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    plain = torch.zeros_like(bake)
    curve = {}
    plain_ms = 0.0
    for it in range(max(frames)):
        ev0.record(stream)
        tracer.execute(enc, plain, it, reset=(it == 0))
        ev1.record(stream)
        torch.cuda.synchronize()
        plain_ms += ev0.elapsed_time(ev1)
        if (it + 1) & it == 0:
            curve[it + 1] = rmse(plain)
    plain_ms /= max(frames)
    del plain
    forms = learned_forms_vs_plain(device, scene, tracer, rmse, curve, plain_ms, frames)
    return {"bake_frames": bake_frames, "plain_ms_per_frame": plain_ms,
            "plain_rmse": {str(k): v for k, v in curve.items()}, "forms": forms}


def bench_learned(args) -> int:
    """The learned radiance cache against the plain tracer at equal RMSE."""
    from nan_amd.device import Device

    device = Device(0)
    t0 = time.perf_counter()
    res = learned_measurement(device)
    log(f"learned measurement {time.perf_counter() - t0:.1f} s")
    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": "effective converged texel-path samples/s (learned, 256 frames)",
        "n_gpus": 1,
        "steps": 256,
        "warmup": 0,
        "ms_per_step": res["learned"]["256"]["learned_ms_per_frame"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 MFMA / f32",
        "data": "synthetic",
        "config": {"workload": res["workload"], "parallelism": "single GPU"},
        "learned_vs_plain": res,
        "roofline": {k: v for k, v in res["mlp_roofline"].items()},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    return 0


COMBINED_LAYOUT = dict(texels_per_unit=2.0, min_resolution=1, max_resolution=4, max_total_texels=16_777_216)


def bench_combined(args) -> int:
    """The combined full-size leg alone (see combined_leg), one JSON line."""
    from nan_amd.device import Device

    res = combined_leg(args, Device(0))
    out = {"metric": METRIC, "value": res["plain_samples_per_s"], "unit": "texel-path samples/s", "n_gpus": 1,
           "steps": res["frames_timed"], "warmup": args.warmup, "ms_per_step": res["latency_ms"]["mean"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": res["workload"], "parallelism": "single GPU"}, "combined": res,
           "roofline": res["mlp_roofline"], "cpu_baseline": None}
    print(json.dumps(out), flush=True)
    return 0


def combined_leg(args, device) -> dict:
    """Combined full-size run: the synthetic generator scaled 3.3x (4.0M
    triangles; 45% of object materials double-sided, terrain single-sided, so
    20.5% of the 16.7M slots are back slots), traced by the batched persistent
    wavefront with every frame a replayed hipGraph (per-frame latency
    percentiles), a 3840x2160 resolve, and the learned forms on the same
    scene against a 4096-frame plain bake."""
    import numpy as np
    import torch

    from nan_amd import _native
    from nan_amd.autotune import autotune
    from nan_amd.batch import SceneBatch
    from nan_amd.passes import MeshColorsResolve
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scenes import synthetic_scene

    dev = device.torch_device
    t0 = time.perf_counter()
    node = synthetic_scene(0, 3.3, double_sided=0.45, single_sided_terrain=True)
    batch = SceneBatch(device, [node], [COMBINED_LAYOUT], [3])
    layout, scene, tracer = batch.layouts[0], batch.scenes[0], batch.tracers[0]
    slots = layout.total_payload_count
    log(f"combined scene {node.triangle_count} tris, {slots} slots ({slots - layout.total_surface_texels} back), "
        f"built in {time.perf_counter() - t0:.1f} s")
    enc = device.create_command_encoder()
    stream = torch.cuda.current_stream(dev)
    tune = autotune(batch, enc, space={"blocks_per_cu": (0, 4), "lds_stack": (16, 32), "refill_steps": (16, 32)},
                    frames=1)
    log(f"autotuned {tune['best']}")
    for p in batch.payloads:
        p.zero_()
    graph = batch.capture(enc, 0, reset=True)  # the first frame itself is a replay
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    frames = max(args.steps, 64)  # enough frames for the latency percentiles
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(frames)]
    for a, b in evs:
        a.record(stream)
        graph.replay()
        b.record(stream)
    torch.cuda.synchronize()
    lat = np.array([a.elapsed_time(b) for a, b in evs])
    batch.clear_device_frame()
    counters = torch.zeros(8, dtype=torch.int64, device=dev)
    scratch_iter = args.warmup + frames
    batch.execute(enc, scratch_iter, counters=counters)  # counts of one frame (payload keeps accumulating)
    torch.cuda.synchronize()
    c = counters.cpu().numpy().astype(np.int64)
    ext, shd = int(c[_native.COUNTER_EXTEND_RAYS]), int(c[_native.COUNTER_SHADOW_RAYS])
    alg = (BYTES_PER_NODE * int(c[_native.COUNTER_NODE_VISITS] + c[_native.COUNTER_SHADOW_NODE_VISITS])
           + BYTES_PER_TRI * int(c[_native.COUNTER_TRI_TESTS] + c[_native.COUNTER_SHADOW_TRI_TESTS])
           + (BYTES_PER_RAY + BYTES_PER_HIT) * ext + BYTES_PER_SHADOW_RAY * shd
           + BYTES_PER_SLOT_UPDATE * int(c[_native.COUNTER_SLOT_UPDATES]) + BYTES_PER_SHADED_HIT * ext)
    mean_ms = float(lat.mean())

    # 3840x2160 resolve of the accumulated cache
    resolve = MeshColorsResolve(device, scene, layout, tracer)
    image = torch.zeros((2160, 3840, 4), dtype=torch.float32, device=dev)
    resolve.execute(enc, batch.payloads[0], image)
    r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    r0.record(stream)
    for _ in range(5):
        resolve.execute(enc, batch.payloads[0], image)
    r1.record(stream)
    torch.cuda.synchronize()
    resolve_ms = r0.elapsed_time(r1) / 5

    # learned form against a plain bake, on the scene's own tracer
    bake = torch.zeros((slots, 4), dtype=torch.float32, device=dev)
    tb = time.perf_counter()
    for it in range(args.bake_frames):
        tracer.execute(enc, bake, 1_000_000 + it, reset=(it == 0))
        if it % 256 == 255:
            torch.cuda.synchronize()
            log(f"bake {it + 1}/{args.bake_frames} ({time.perf_counter() - tb:.0f} s)")
    torch.cuda.synchronize()
    bk = bake[:, :3].double()

    def rmse(p):
        return float(torch.sqrt(torch.mean((p[:, :3].double() - bk) ** 2)))

    plain = torch.zeros_like(bake)
    curve = {}
    for it in range(64):
        tracer.execute(enc, plain, it, reset=(it == 0))
        if (it + 1) & it == 0:
            curve[it + 1] = rmse(plain)
    forms = learned_forms_vs_plain(device, scene, tracer, rmse, curve, mean_ms, graph=True)
    # one captured graph holding the batched plain frame and the learned
    # frame (the specified one-bounce form) of the same scene, each advancing
    # its own device frame per replay
    n_train, mb = LEARNED_FORMS["one_bounce"][1:]
    cache = RadianceCache.for_scene(device, scene, cache_config(scene, tracer.layout, mb))
    learned = torch.zeros_like(bake)
    tracer.reserve_learned(cache, n_train)
    batch.set_device_frame(2_000_000, True)
    tracer.set_device_frame(0, True)
    torch.cuda.synchronize()
    both = torch.cuda.CUDAGraph()
    with torch.cuda.graph(both):
        batch.execute(enc, 0)
        tracer.execute_learned(enc, learned, 0, cache, train_samples=n_train)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(frames)]
    for a, b in evs:
        a.record(stream)
        both.replay()
        b.record(stream)
    torch.cuda.synchronize()
    lat_both = np.array([a.elapsed_time(b) for a, b in evs])
    batch.clear_device_frame()
    tracer.clear_device_frame()
    del both
    one_graph = {"frames": frames, "learned_rmse": rmse(learned),
                 "latency_ms": {"p50": float(np.percentile(lat_both, 50)), "p99": float(np.percentile(lat_both, 99)),
                                "mean": float(lat_both.mean())}}
    log(f"plain + learned in one graph: p50 {one_graph['latency_ms']['p50']:.2f} ms")
    q = slots  # one cache query per texel sample
    rng = np.random.default_rng(3)
    lo, hi = scene.bounds()
    pos = torch.from_numpy(rng.uniform(lo, hi, size=(q, 3)).astype(np.float32)).to(dev)
    nrm = torch.nn.functional.normalize(torch.randn((q, 3), device=dev), dim=1).contiguous()
    alb = torch.rand((q, 3), device=dev)
    mlp_ms = cache.time_forward(pos, nrm, alb, repeats=5, encoder=enc)
    tflops = MLP_FLOPS_PER_QUERY * q / (mlp_ms * 1e-3) / 1e12
    return {
        "workload": ("combined: synthetic generator scaled 3.3x (seed 0, 45% of object materials double-sided, "
                     "terrain single-sided), layout texels_per_unit 2, res 1-4, budget 16,777,216; batched "
                     "persistent wavefront (batch of one scene), every frame a replayed hipGraph; 1 spp, "
                     "3 bounces; resolve 3840x2160; learned forms on the same scene"),
        "triangles": node.triangle_count,
        "payload_slots": slots,
        "back_slots": slots - layout.total_surface_texels,
        "layout": COMBINED_LAYOUT,
        "frames_timed": frames,
        "plain_samples_per_s": slots / (mean_ms * 1e-3),
        "latency_ms": {"mean": mean_ms, "p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "min": float(lat.min()), "max": float(lat.max())},
        "mrays_per_s": (ext + shd) / (mean_ms * 1e-3) / 1e6,
        "rays_per_step": {"extend": ext, "shadow": shd},
        "algorithmic_gb_per_s": alg / (mean_ms * 1e-3) / 1e9,
        "autotune": tune["best"],
        "resolve_3840x2160_ms": resolve_ms,
        "learned_vs_plain": {"bake_frames": args.bake_frames, "plain_rmse": {str(k): v for k, v in curve.items()},
                             "forms": forms},
        "plain_and_learned_one_graph": one_graph,
        "mlp_roofline": {"bound": "mfma", "kernel": "nrc_mlp_fwd<false>", "queries": q, "avg_launch_ms": mlp_ms,
                         "achieved": tflops, "peak": MFMA_F16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": tflops / MFMA_F16_PEAK_TFLOPS, "instruction": MFMA_INSTRUCTION,
                         "instruction_peak": MFMA_K16_PEAK_TFLOPS,
                         "instruction_frac": tflops / MFMA_K16_PEAK_TFLOPS},
    }


def batch_leg(args, device) -> dict:
    """64 seeded scenes (1k-200k triangles, 16k-2M slot budgets, Russian
    roulette paths of 1-8 bounces) traced by one set of wavefront launches
    captured in a hipGraph, after the autotuner picked the traversal launch
    shape; ``args.steps`` replayed frames timed with HIP events."""
    import torch

    from nan_amd.autotune import autotune
    from nan_amd.batch import SceneBatch, batch_workload

    t0 = time.perf_counter()
    nodes, kws, bounces = batch_workload(0, 64)
    batch = SceneBatch(device, nodes, kws, bounces)
    log(f"batch of {len(nodes)} scenes, {sum(batch.triangle_counts)} tris, {batch.total_slots} slots "
        f"({time.perf_counter() - t0:.1f} s)")
    enc = device.create_command_encoder()
    tuned = autotune(batch, enc, frames=1, log=log)
    # frame 0 eagerly, then replay a captured frame
    batch.execute(enc, 0, reset=True)
    for it in range(1, 1 + args.warmup):
        batch.execute(enc, it)
    torch.cuda.synchronize()
    graph = batch.capture(enc, 1 + args.warmup)
    stream = torch.cuda.current_stream()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        graph.replay()
    end.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(end)
    batch.clear_device_frame()
    slots = batch.total_slots
    return {
        "workload": "64-scene batch (seed 0): 1k-200k triangles and 16k-2M slot budgets log-uniform, "
                    "Russian roulette paths of 1-8 bounces, one hipGraph per frame",
        "value": slots * args.steps / (total_ms * 1e-3),
        "unit": "texel-path samples/s",
        "ms_per_step": total_ms / args.steps,
        "steps": args.steps,
        "scenes": len(nodes),
        "russian_roulette": True,
        "triangles": int(sum(batch.triangle_counts)),
        "payload_slots": int(slots),
        "bounces": bounces,
        "tuning": tuned["best"],
        "autotune": tuned["results"],
    }


def bench_batch(args) -> int:
    """The 64-scene batch alone (batch_leg)."""
    from nan_amd.device import Device

    res = batch_leg(args, Device(0))
    out = {
        "metric": METRIC,
        "value": res["value"],
        "unit": res["unit"],
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {k: res[k] for k in ("workload", "scenes", "russian_roulette", "triangles", "payload_slots",
                                       "bounces", "tuning")},
        "autotune": res["autotune"],
    }
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["synthetic", "cornell", "batch", "learned", "combined"],
                    default="synthetic")
    ap.add_argument("--bake-frames", type=int, default=4096, help="plain bake length of --workload combined")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-cache comparison")
    ap.add_argument("--scale-bake-frames", type=int, default=4096,
                    help="plain bake length of the learned comparison on the workload's own scene")
    ap.add_argument("--no-combined", action="store_true", help="skip the combined full-size leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch leg")
    ap.add_argument("--mode", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (counter calibration)")
    ap.add_argument("--calibrate-valu", action="store_true",
                    help="only run the VALU-bound FMA kernel (issue-rate calibration of SQ_INSTS_VALU)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch

        n = 1 << 28  # 2^28 float32 = 1 GiB
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "calibrate_copy_bytes_written": 4 * n, "copies": 3}))
        return 0

    if args.calibrate_valu:
        import ctypes

        from nan_amd import _native

        ms, insts = ctypes.c_double(), ctypes.c_double()
        _native.check(_native.lib().nan_calibrate_valu(20000, ctypes.byref(ms), ctypes.byref(insts)),
                      "nan_calibrate_valu")
        print(json.dumps({"calibrate_valu_ms": ms.value, "fma_wave_insts": insts.value,
                          "wave_insts_per_s": insts.value / (ms.value * 1e-3)}))
        return 0

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; the path shards by texel range with no "
                                   "exchange during tracing, and multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "learned":
        return bench_learned(args)
    if args.workload == "combined":
        return bench_combined(args)

    return bench_default(args)


def plain_leg(args, device) -> tuple[dict, tuple]:
    """The plain tracer on the workload's scene: K timed iterations of the
    production path, a profiled replay (every launch timed alone, one-stream
    order) for the stage times and the roofline's launch duration, and a
    counted replay for the work counts.  Returns the plain block and
    (node, layout, scene, tracer, stages, shadow_rays)."""
    import numpy as np
    import torch

    from nan_amd import _native
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    t0 = time.perf_counter()
    node, layout, layout_kw, desc = build_workload(args.workload)
    log(f"scene {node.triangle_count} tris, {len(node.instances)} instances; layout "
        f"{layout.total_surface_texels} surface + {layout.total_payload_count - layout.total_surface_texels} back "
        f"slots ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    scene = Scene(device, node)
    log(f"BVH: {scene.info.node_count} nodes, {scene.info.leaf_count} leaves, max stack {scene.info.max_stack}, "
        f"build {scene.info.build_ms:.0f} ms, upload total {time.perf_counter() - t0:.1f} s")
    tracer = TextureSpacePathTracer(device, scene, layout, samples_per_texel=1, max_bounces=3, mode=args.mode)
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    enc = device.create_command_encoder()
    stream = torch.cuda.current_stream(device.torch_device)

    it = 0
    for _ in range(args.warmup):
        tracer.execute(enc, payload, it, reset=(it == 0))
        it += 1
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} iterations done ({args.mode})")

    first_timed = it
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        tracer.execute(enc, payload, it, reset=False)
        ends[k].record(stream)
        it += 1
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    total_ms = starts[0].elapsed_time(ends[-1])
    # Stage profile: the same frames again with HIP events around every
    # launch.  Profiling runs the sweeps in one-stream order (the timed frames
    # trace the double-sided texels' paths on a second stream beside the
    # rest), so every launch is timed alone.
    scratch = payload.clone()
    tracer.profile(True)
    p0 = torch.cuda.Event(enable_timing=True)
    p1 = torch.cuda.Event(enable_timing=True)
    p0.record(stream)
    for k in range(args.steps):
        tracer.execute(enc, scratch, first_timed + k, reset=False)
    p1.record(stream)
    torch.cuda.synchronize()
    stages = tracer.stage_times()
    tracer.profile(False)
    profiled_ms = p0.elapsed_time(p1)
    del scratch
    log(f"timed {args.steps} iterations: {total_ms:.2f} ms (events), {wall * 1e3:.2f} ms (wall); profiled replay "
        f"{profiled_ms:.2f} ms, stages " + ", ".join(f"{k} {v[0] / args.steps:.2f} ms" for k, v in stages.items()))

    # counted replay of the same iterations on a scratch copy (counts depend
    # only on the iteration, not on the accumulated values)
    scratch = payload.clone()
    counters = torch.zeros(8, dtype=torch.int64, device=device.torch_device)
    for k in range(args.steps):
        tracer.execute(enc, scratch, first_timed + k, reset=False, counters=counters)
    torch.cuda.synchronize()
    c = counters.cpu().numpy().astype(np.int64) // args.steps
    del scratch
    ext, shd = int(c[_native.COUNTER_EXTEND_RAYS]), int(c[_native.COUNTER_SHADOW_RAYS])
    nodes, tris = int(c[_native.COUNTER_NODE_VISITS]), int(c[_native.COUNTER_TRI_TESTS])
    snodes, stris = int(c[_native.COUNTER_SHADOW_NODE_VISITS]), int(c[_native.COUNTER_SHADOW_TRI_TESTS])
    slots = int(c[_native.COUNTER_SLOT_UPDATES])
    front = layout.total_surface_texels
    value = slots * args.steps / (total_ms * 1e-3)
    mrays = (ext + shd) * args.steps / (total_ms * 1e-3) / 1e6

    # algorithmic bytes per frame of each stage
    stage_bytes = {
        "extend": nodes * BYTES_PER_NODE + tris * BYTES_PER_TRI + ext * (BYTES_PER_RAY + BYTES_PER_HIT),
        "shadow": snodes * BYTES_PER_NODE + stris * BYTES_PER_TRI + shd * BYTES_PER_SHADOW_RAY,
        "megakernel": ((nodes + snodes) * BYTES_PER_NODE + (tris + stris) * BYTES_PER_TRI
                       + slots * BYTES_PER_SLOT_UPDATE + front * BYTES_PER_TEXEL_SETUP
                       + (slots + ext) * BYTES_PER_SHADED_HIT),
    }
    dominant = max((k for k in stages if k in stage_bytes), key=lambda k: stages[k][0])
    dom_ms, dom_launches = stages[dominant]
    bytes_per_launch = stage_bytes[dominant] * args.steps / dom_launches
    avg_launch_ms = dom_ms / dom_launches
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    kernel_name = {"extend": "wf_extend", "shadow": "wf_shadow", "megakernel": "tst_megakernel"}[dominant]

    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None,
        "kernel": kernel_name,
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "avg_launch_ms": avg_launch_ms,
        "launches_per_step": dom_launches / args.steps,
    }
    # HBM traffic, L2 and issue counters of this kernel on this workload, from
    # the committed rocprofv3 summary (copy-calibrated bytes, FMA-calibrated
    # VALU issue); absent when the summary is for another kernel version or
    # workload.  The algorithmic bytes count every node and triangle read,
    # most of them served by L2 and the MALL, so `frac` can reach 1 while HBM
    # runs far below it: `bound` names the resource the counters show closest
    # to saturation.
    prof = ROOT / "profiles" / f"traffic_{kernel_name}_{args.workload}.json"
    if prof.exists():
        rec = json.loads(prof.read_text())
        if str(layout.total_payload_count) in rec.get("workload", "") and rec.get("kernel_version") == KERNEL_VERSION:
            launch_s = avg_launch_ms * 1e-3
            roofline.update({
                "bound": rec.get("bound", "hbm"),
                "traffic": rec["traffic_bytes_per_launch"],
                "hbm_gbs": rec["traffic_bytes_per_launch"] / launch_s / 1e9,
                "hbm_frac": rec["traffic_bytes_per_launch"] / launch_s / 1e9 / HBM_PEAK_GBS,
                "l2_hit_rate": rec.get("l2_hit_rate"),
                "l2_read_requests_per_s": (rec["l2_read_requests_per_launch"] / launch_s
                                           if "l2_read_requests_per_launch" in rec else None),
                "valu_issue_frac": rec.get("valu_issue_frac"),
                "valu_dual_issue_share": rec.get("valu_dual_issue_share"),
                "wave_wait_frac": rec.get("wait_frac"),
                "profile": rec.get("source"),
            })
    plain = {
        "workload": desc,
        "value": value,
        "unit": "texel-path samples/s",
        "ms_per_step": total_ms / args.steps,
        "steps": args.steps,
        "warmup": args.warmup,
        "triangles": node.triangle_count,
        "instances": len(node.instances),
        "payload_slots": layout.total_payload_count,
        "surface_texels": layout.total_surface_texels,
        "layout": layout_kw,
        "samples_per_texel": 1,
        "max_bounces": 3,
        "tracer": args.mode,
        "mrays_per_s": mrays,
        "rays_per_step": {"extend": ext, "shadow": shd},
        "traversal_per_step": {"extend_node_visits": nodes, "extend_tri_tests": tris,
                               "shadow_node_visits": snodes, "shadow_tri_tests": stris},
        "stage_ms_per_step": {k: v[0] / args.steps for k, v in stages.items()},
        "profiled_ms_per_step": profiled_ms / args.steps,
        "roofline": roofline,
    }
    return plain, (node, layout, scene, tracer, stages, shd)


def host_cpu() -> dict:
    """The host's CPU model and counts (the CPU baseline's threads are the
    process's CPU affinity, at most 16: one GPU's share of the host)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "os_cpu_count": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def bench_default(args) -> int:
    """The default run: the metric on its configuration (the learned form
    against the plain tracer on the Cornell box, learned_vs_plain), the plain
    tracer at scale with its roofline, both learned forms at scale, the
    combined full-size leg and the CPU baselines."""
    from nan_amd.device import Device

    device = Device(0)
    plain, (node, layout, scene, tracer, stages, shd) = plain_leg(args, device)
    out_learned = at_scale = combined = None
    if not args.no_learned:
        t0 = time.perf_counter()
        out_learned = learned_measurement(device, steps=args.steps, warmup=args.warmup)
        log(f"learned measurement {time.perf_counter() - t0:.1f} s")
        if args.workload == "synthetic":
            t0 = time.perf_counter()
            at_scale = learned_at_scale(device, scene, layout, tracer, bake_frames=args.scale_bake_frames)
            log(f"learned at scale {time.perf_counter() - t0:.1f} s")
    batch = None
    if not args.no_batch and args.workload == "synthetic":
        t0 = time.perf_counter()
        batch = batch_leg(args, device)
        log(f"batch leg {time.perf_counter() - t0:.1f} s: {batch['ms_per_step']:.2f} ms/frame")
    if not args.no_combined and args.workload == "synthetic":
        t0 = time.perf_counter()
        combined = combined_leg(args, device)
        log(f"combined leg {time.perf_counter() - t0:.1f} s")
    if out_learned is not None:
        head = {
            "value": out_learned["value"],
            "unit": "effective converged texel-path samples/s",
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": out_learned["timed_ms_per_frame"],
            "dtype": "f32 tracing / f16 MFMA network, f32 accumulate",
            "config": {"workload": "learned vs plain at equal RMSE: " + out_learned["workload"],
                       "frames": out_learned["value_frames"], "bake_frames": out_learned["bake_frames"],
                       "payload_slots": out_learned["payload_slots"], "parallelism": "single GPU"},
            "value_vs_plain": out_learned["speedup_vs_plain"],
        }
    else:
        head = {
            "value": plain["value"], "unit": plain["unit"], "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": plain["ms_per_step"], "dtype": "f32",
            "config": {"workload": plain["workload"], "parallelism": "single GPU"},
        }
    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": head["unit"],
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": head["warmup"],
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head["dtype"],
        "data": "synthetic",
        "config": head["config"],
    }
    if "value_vs_plain" in head:
        out["value_vs_plain"] = head["value_vs_plain"]
        out["learned_vs_plain"] = out_learned
    out["plain_samples_per_s"] = plain["value"]
    out["plain_at_scale"] = {k: v for k, v in plain.items() if k != "roofline"}
    out["roofline"] = plain["roofline"]
    if at_scale is not None:
        out["learned_vs_plain_at_scale"] = at_scale
    if batch is not None:
        out["batch"] = batch
    if combined is not None:
        out["combined"] = combined
    out["cpu_baseline"] = None
    if not args.no_cpu_baseline:
        threads = min(len(os.sched_getaffinity(0)), 16)
        log(f"cpu baseline on {threads} threads")
        out["cpu_baseline"] = cpu_baseline(node, layout, threads, args.cpu_seconds)
        out["cpu_baseline"]["host"] = host_cpu()
        ref = tinybvh_baseline(node, threads)
        if ref is not None:
            ref["gpu_shadow_mrays_per_s"] = shd / (stages["shadow"][0] / args.steps * 1e-3) / 1e6
            ref["host"] = host_cpu()
            out["cpu_baseline_reference_visibility"] = ref
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
