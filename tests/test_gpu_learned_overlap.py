"""The learned form with overlapped training (NAN_LEARNED_OVERLAP): a call
copies the cache's parameters for its queries, trains the cache on the
previous call's training records on a second stream while it traces and
queries the copy, and keeps its own records for the next call.  Checked
against the cache trained explicitly on those records, against calls that
do not train (what the queries read), for convergence on the Cornell box,
and under hipGraph replay against eager calls."""

from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg: str) -> None:
    print(f"[gpu-overlap] {msg}", flush=True)
    sys.stdout.flush()


def _setup(device, cornell):
    from nan_amd.radiance_cache import RadianceCacheConfig
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    node, layout, _ = cornell
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="megakernel")
    lo, hi = sc.bounds()
    cfg = RadianceCacheConfig.for_texel_density(float(np.max(hi - lo)), layout.texels_per_unit, minibatches=2)
    return layout, sc, tracer, cfg


def _same_updates(dg: np.ndarray, de: np.ndarray) -> None:
    """Equal up to the hash-grid atomics' rounding order: 99% of the entries
    within 2% or 2e-4, update norms within 2%."""
    close = np.isclose(dg, de, rtol=2e-2, atol=2e-4)
    assert close.mean() > 0.99, close.mean()
    ratio = float(np.linalg.norm(dg) / np.linalg.norm(de))
    assert abs(ratio - 1.0) < 0.02, ratio


def test_overlapped_call_trains_on_the_previous_calls_records(device, cornell):
    """The first overlapped call's steps see no valid record and leave the
    parameters unchanged; the second call trains on the first call's records.
    Its twin: a cache that takes the same zero steps (an overlapped call on a
    tracer of its own), then trains on iteration 0's records without overlap
    (so both train on the same records with the same valid flags at the same
    Adam step counts)."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    layout, sc, tracer, cfg = _setup(device, cornell)
    n = 16384
    caches = [RadianceCache.for_scene(device, sc, cfg) for _ in range(2)]
    w0, t0 = caches[0].get_params()
    enc = device.create_command_encoder()
    out = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    _say("first overlapped call")
    tracer.execute_learned(enc, out, 0, caches[0], reset=True, train_samples=n, overlap=True)
    w1, t1 = caches[0].get_params()
    np.testing.assert_array_equal(w1, w0)
    np.testing.assert_array_equal(t1, t0)
    _say("second overlapped call vs the twin")
    tracer.execute_learned(enc, out, 1, caches[0], train_samples=n, overlap=True)
    twin_tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="megakernel")
    twin = torch.zeros_like(out)
    twin_tracer.execute_learned(enc, twin, 7, caches[1], reset=True, train_samples=n, overlap=True)
    twin_tracer.execute_learned(enc, twin, 0, caches[1], reset=True, train_samples=n)
    wa, ta = caches[0].get_params()
    wb, tb = caches[1].get_params()
    assert np.abs(wb - w0).max() > 1e-3
    _same_updates(wa - w0, wb - w0)
    _same_updates((ta - t0).ravel(), (tb - t0).ravel())
    v = out.cpu().numpy()
    assert np.isfinite(v).all() and (v.view(np.uint32)[:, 3] == 2).all()


def test_overlapped_queries_read_the_parameters_the_previous_call_left(device, cornell):
    """The queries of an overlapped call read the parameters copied at the
    call's start, before its own steps.  The first call's step sees no valid
    record, so the first two calls both query the untrained cache: their
    payload equals that of two calls that do not train, bit for bit, although
    the second call's step did change the cache."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    layout, sc, tracer, cfg = _setup(device, cornell)
    n = 16384
    caches = [RadianceCache.for_scene(device, sc, cfg) for _ in range(2)]
    w0, _ = caches[0].get_params()
    enc = device.create_command_encoder()
    out = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    twin = torch.zeros_like(out)
    twin_tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="megakernel")
    _say("two overlapped calls vs two calls without training")
    for it in range(2):
        tracer.execute_learned(enc, out, it, caches[0], reset=(it == 0), train_samples=n, overlap=True)
        twin_tracer.execute_learned(enc, twin, it, caches[1], reset=(it == 0), train_samples=0)
    w2, _ = caches[0].get_params()
    assert np.abs(w2 - w0).max() > 1e-3
    a, b = out.cpu().numpy(), twin.cpu().numpy()
    assert (a.view(np.uint32)[:, 3] == 2).all()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_overlapped_training_converges_on_the_cornell_box(device, cornell):
    """After 64 overlapped frames (eight Adam steps of 8,192 suffixes each)
    the learned form is unbiased in the mean (within 5% of a 512-frame plain
    bake) and within 1.5x of the plain tracer's RMSE."""
    from nan_amd.radiance_cache import RadianceCache, RadianceCacheConfig

    layout, sc, tracer, cfg = _setup(device, cornell)
    cfg = RadianceCacheConfig(**{**cfg.__dict__, "minibatches": 8})
    enc = device.create_command_encoder()
    slots = layout.total_payload_count
    bake = torch.zeros((slots, 4), dtype=torch.float32, device=device.torch_device)
    for it in range(512):
        tracer.execute(enc, bake, 100_000 + it, reset=(it == 0))
    cache = RadianceCache.for_scene(device, sc, cfg)
    learned = torch.zeros_like(bake)
    plain = torch.zeros_like(bake)
    for it in range(64):
        tracer.execute_learned(enc, learned, it, cache, reset=(it == 0), overlap=True)
        tracer.execute(enc, plain, it, reset=(it == 0))
    b = bake[:, :3].cpu().numpy().astype(np.float64)
    lq = learned[:, :3].cpu().numpy().astype(np.float64)
    pq = plain[:, :3].cpu().numpy().astype(np.float64)
    assert np.isfinite(lq).all()
    rmse_l = np.sqrt(np.mean((lq - b) ** 2))
    rmse_p = np.sqrt(np.mean((pq - b) ** 2))
    _say(f"64 frames: RMSE learned {rmse_l:.4f} plain {rmse_p:.4f}; means {lq.mean():.4f} {b.mean():.4f}")
    assert abs(lq.mean() - b.mean()) < 0.05 * b.mean()
    assert rmse_l < 1.5 * rmse_p


def test_overlapped_iterations_replayed_from_one_capture(device, cornell):
    """Device-frame mode: one captured overlapped iteration (its fork and
    join inside the graph) replayed three times equals three eager calls on a
    twin tracer and cache, to the tolerance of the atomics' rounding order."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    layout, sc, _, cfg = _setup(device, cornell)
    # device-frame mode is a wavefront-mode feature; overlapped calls run the
    # per-texel kernels in either mode
    tracer, eager = (TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="wavefront")
                     for _ in range(2))
    n = 8192
    caches = [RadianceCache.for_scene(device, sc, cfg) for _ in range(2)]
    w0, _ = caches[0].get_params()
    dev = device.torch_device
    enc = device.create_command_encoder()
    outs = [torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev) for _ in range(2)]
    _say("capture an overlapped iteration, replay three times")
    g = tracer.capture_learned(enc, outs[0], caches[0], 0, reset=True, train_samples=n, overlap=True)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
    tracer.clear_device_frame()
    for it in range(3):
        eager.execute_learned(enc, outs[1], it, caches[1], reset=(it == 0), train_samples=n, overlap=True)
    torch.cuda.synchronize()
    wg, _ = caches[0].get_params()
    we, _ = caches[1].get_params()
    assert np.abs(we - w0).max() > 1e-3
    _same_updates(wg - w0, we - w0)
    a, b = outs[0].cpu().numpy(), outs[1].cpu().numpy()
    assert (a.view(np.uint32)[:, 3] == 3).all() and (b.view(np.uint32)[:, 3] == 3).all()
    rel = np.abs(a[:, :3] - b[:, :3]).mean() / np.abs(b[:, :3]).mean()
    assert rel < 1e-2, rel
