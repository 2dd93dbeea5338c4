"""Graph capture of first frames, the accumulate edge cases and long paths
on the GPU.

* A tracer's very first execute, and the first execute after a launch-shape
  change, captured into a hipGraph and replayed, equal eager runs bit for
  bit: the working set is allocated at creation and by tune, never by an
  execute.
* The first learned iteration captured after reserve_learned and replayed
  twice trains the cache like two eager iterations: Adam's step count lives
  on the device, so the second replay takes step 2's bias corrections.
* Count saturation at 65535 and the reset of a non-zero slot, against the
  oracle (texture_space_path_tracer.slang:119-138).
* 16 samples per texel with 8 bounces (more queue counters than one frame's
  budget, reused per sample): wavefront == megakernel == oracle.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from nan_amd.mesh_colors import MeshColorsLayout  # noqa: E402
from nan_amd.scene_arrays import flatten_scene  # noqa: E402
from nan_amd.scenes import batch_scene, cornell_box, synthetic_scene  # noqa: E402


def _say(msg: str) -> None:
    print(f"[gpu-capture] {msg}", flush=True)
    sys.stdout.flush()


def _tracer(device, node, mode="wavefront", spp=1, bounces=3, **kw):
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    layout = MeshColorsLayout.build(node, **kw)
    sc = Scene(device, node)
    return layout, sc, TextureSpacePathTracer(device, sc, layout, mode=mode, samples_per_texel=spp,
                                              max_bounces=bounces)


def _bits_equal(a: torch.Tensor, b: torch.Tensor, tag: str) -> None:
    x = a.cpu().numpy().view(np.uint32)
    y = b.cpu().numpy().view(np.uint32)
    bad = np.nonzero(np.any(x != y, axis=1))[0]
    assert len(bad) == 0, f"{tag}: {len(bad)} of {len(x)} slots differ, first {bad[0]}"


def _same_updates(dg: np.ndarray, de: np.ndarray) -> None:
    """Weight updates of a replayed and an eager run: equal up to the
    atomics' rounding order (99% of weights within 2% or 2e-4, update norms
    within 2%)."""
    close = np.isclose(dg, de, rtol=2e-2, atol=2e-4)
    assert close.mean() > 0.99, close.mean()
    ratio = float(np.linalg.norm(dg) / np.linalg.norm(de))
    assert abs(ratio - 1.0) < 0.02, ratio


def test_first_execute_and_first_after_tune_captured(device):
    # 1/16 of the texels double-sided: the wavefront forks its second stream inside the capture
    node = synthetic_scene(3, terrain_quads=48, box_grid=8, spheres=24, extent=20.0)
    kw = dict(texels_per_unit=8.0, min_resolution=2, max_resolution=8, max_total_texels=4_000_000)
    layout, sc, tracer = _tracer(device, node, **kw)
    _, _, eager = _tracer(device, node, **kw)
    assert layout.total_payload_count > layout.total_surface_texels
    dev = device.torch_device
    enc = device.create_command_encoder()
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
    ref = torch.zeros_like(payload)
    _say("capturing the tracer's first execute")
    torch.cuda.synchronize()
    g0 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g0):
        tracer.execute(enc, payload, 0, reset=True)
    g0.replay()
    eager.execute(enc, ref, 0, reset=True)
    torch.cuda.synchronize()
    _bits_equal(payload, ref, "first execute, captured")
    _say("tune, then capture the next frame straight away")
    tracer.tune(2, 24, 8)
    g1 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g1):
        tracer.execute(enc, payload, 1)
    g1.replay()
    eager.execute(enc, ref, 1)
    torch.cuda.synchronize()
    _bits_equal(payload, ref, "first execute after tune, captured")
    assert torch.isfinite(payload).all() and payload[:, :3].max() > 0


def test_first_learned_iteration_captured_and_replayed(device):
    from nan_amd.radiance_cache import RadianceCache, RadianceCacheConfig

    node = cornell_box()
    layout, sc, tracer = _tracer(device, node)
    dev = device.torch_device
    enc = device.create_command_encoder()
    train = 16384
    caches = [RadianceCache.for_scene(device, sc, RadianceCacheConfig()) for _ in range(2)]
    w0, _ = caches[0].get_params()
    w0b, _ = caches[1].get_params()
    np.testing.assert_array_equal(w0, w0b)
    payloads = [torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev) for _ in range(2)]
    _say("learned: reserve, capture the first iteration, replay twice")
    tracer.reserve_learned(caches[0], train)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        tracer.execute_learned(enc, payloads[0], 0, caches[0], reset=True, train_samples=train)
    for _ in range(2):
        g.replay()
        torch.cuda.synchronize()
    _say("learned: two eager iterations on a twin cache")
    for _ in range(2):
        tracer.execute_learned(enc, payloads[1], 0, caches[1], reset=True, train_samples=train)
    torch.cuda.synchronize()
    wg, _ = caches[0].get_params()
    we, _ = caches[1].get_params()
    # the hash-grid gradients are summed with atomics (order-dependent FP16
    # rounding), so the twins agree closely, not bit for bit (weights whose
    # gradient is near zero take Adam's sign-like step either way); a replay
    # stuck at step 1's bias corrections would move every weight ~1.9x as far
    # in step 2
    dg, de = wg - w0, we - w0
    assert np.abs(de).max() > 1e-3
    _same_updates(dg, de)
    a = payloads[0][:, :3].double()
    b = payloads[1][:, :3].double()
    assert torch.isfinite(a).all()
    rel = float((a - b).abs().mean() / b.abs().mean())
    assert rel < 1e-2, rel


def test_count_saturation_and_reset_of_nonzero_slots(device):
    node = batch_scene(21, 3000)
    kw = dict(texels_per_unit=6.0, min_resolution=1, max_resolution=8, max_total_texels=2_000_000)
    for mode in ("wavefront", "megakernel"):
        layout, sc, tracer = _tracer(device, node, mode=mode, **kw)
        rng = np.random.default_rng(17)
        n = layout.total_payload_count
        start = np.zeros((n, 4), dtype=np.float32)
        start[:, :3] = rng.uniform(0.0, 5.0, size=(n, 3)).astype(np.float32)
        counts = rng.choice(np.array([65533, 65534, 65535, 1, 77], dtype=np.uint32), size=n)
        start.view(np.uint32)[:, 3] = counts
        orc = oracle.OracleScene(flatten_scene(node))
        enc = device.create_command_encoder()
        for reset in (False, True):
            _say(f"{mode}: iteration 5 over non-zero slots, reset={reset}")
            payload = torch.from_numpy(start.copy()).to(device.torch_device)
            tracer.execute(enc, payload, 5, reset=reset)
            ref = start.copy()
            orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, 5, 1, 3, reset,
                             threads=oracle.cpu_count())
            got = payload.cpu().numpy()
            bad = np.nonzero(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1))[0]
            assert len(bad) == 0, f"{mode} reset={reset}: {len(bad)} slots differ, first {bad[0]}"
            c = got.view(np.uint32)[:, 3]
            if reset:
                assert (c == 1).all()
            else:
                assert (c[counts >= 65534] == 65535).all() and (c[counts == 77] == 78).all()


def test_sixteen_samples_eight_bounces(device):
    node = batch_scene(22, 1500)
    kw = dict(texels_per_unit=3.0, min_resolution=1, max_resolution=4, max_total_texels=2_000_000)
    outs = []
    for mode in ("wavefront", "megakernel"):
        _say(f"spp 16, 8 bounces, {mode}")
        layout, sc, tracer = _tracer(device, node, mode=mode, spp=16, bounces=8, **kw)
        payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
        enc = device.create_command_encoder()
        for it in range(2):
            tracer.execute(enc, payload, it, reset=(it == 0))
        outs.append(payload.cpu().numpy())
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    orc = oracle.OracleScene(flatten_scene(node))
    ref = np.zeros_like(outs[0])
    for it in range(2):
        orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 16, 8, it == 0,
                         threads=oracle.cpu_count())
    np.testing.assert_array_equal(outs[0].view(np.uint32), ref.view(np.uint32))


def test_learned_iterations_replayed_from_one_capture(device):
    """Device-frame mode: one captured learned iteration replayed three
    times equals three eager iterations (iterations 0, 1, 2 from a reset) on
    a twin cache, to the tolerance of the atomically summed hash-grid
    gradients, on the wavefront queues (a sunlit scene: no emitter light
    sampling)."""
    from nan_amd.radiance_cache import RadianceCache, RadianceCacheConfig

    node = batch_scene(21, 3000)
    kw = dict(texels_per_unit=6.0, min_resolution=1, max_resolution=8, max_total_texels=2_000_000)
    layout, sc, tracer = _tracer(device, node, **kw)
    _, _, eager = _tracer(device, node, **kw)
    dev = device.torch_device
    enc = device.create_command_encoder()
    train = 8192
    caches = [RadianceCache.for_scene(device, sc, RadianceCacheConfig(minibatches=2)) for _ in range(2)]
    w0, _ = caches[0].get_params()
    payloads = [torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev) for _ in range(2)]
    _say("learned: capture in device-frame mode, replay 3 iterations")
    g = tracer.capture_learned(enc, payloads[0], caches[0], 0, reset=True, train_samples=train)
    for _ in range(3):
        g.replay()
    torch.cuda.synchronize()
    tracer.clear_device_frame()
    for it in range(3):
        eager.execute_learned(enc, payloads[1], it, caches[1], reset=(it == 0), train_samples=train)
    torch.cuda.synchronize()
    wg, _ = caches[0].get_params()
    we, _ = caches[1].get_params()
    _same_updates(wg - w0, we - w0)
    a, b = payloads[0].cpu().numpy(), payloads[1].cpu().numpy()
    np.testing.assert_array_equal(a.view(np.uint32)[:, 3], b.view(np.uint32)[:, 3])  # sample counts: 3 iterations
    assert (a.view(np.uint32)[:, 3] == 3).all()
    rel = np.abs(a[:, :3] - b[:, :3]).mean() / np.abs(b[:, :3]).mean()
    assert rel < 1e-2, rel
