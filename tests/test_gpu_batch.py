"""Batched scenes on the GPU: one set of wavefront launches (eager and as a
replayed hipGraph) for several independent scenes must leave every scene's
cache bit-identical to tracing that scene alone, whatever the launch shape."""

from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from nan_amd.scenes import batch_scene, cornell_box  # noqa: E402


def _say(msg: str) -> None:
    print(f"[gpu-batch] {msg}", flush=True)
    sys.stdout.flush()


def _small_batch():
    nodes = [batch_scene(31, 2500), cornell_box(), batch_scene(32, 9000), batch_scene(33, 1200),
             batch_scene(34, 5000)]
    kws = [dict(texels_per_unit=4.0, min_resolution=1, max_resolution=8, max_total_texels=500_000),
           dict(texels_per_unit=4.0, min_resolution=2, max_resolution=16, max_total_texels=500_000),
           dict(texels_per_unit=2.0, min_resolution=1, max_resolution=8, max_total_texels=500_000),
           dict(texels_per_unit=8.0, min_resolution=1, max_resolution=8, max_total_texels=500_000),
           dict(texels_per_unit=3.0, min_resolution=1, max_resolution=8, max_total_texels=500_000)]
    bounces = [1, 8, 3, 5, 2]
    return nodes, kws, bounces


def _single_runs(batch, iterations):
    """Each scene traced alone by its own tracer (same path length)."""
    out = []
    enc = batch.device.create_command_encoder()
    for tracer, layout in zip(batch.tracers, batch.layouts):
        p = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=batch.device.torch_device)
        for it in range(iterations):
            tracer.execute(enc, p, it, reset=(it == 0))
        out.append(p.cpu().numpy())
    return out


def _assert_same(batch, singles, tag):
    for k, (p, ref) in enumerate(zip(batch.payloads, singles)):
        got = p.cpu().numpy()
        bad = np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1)
        assert bad.sum() == 0, f"{tag}: scene {k}: {bad.sum()} of {len(bad)} slots differ"


def test_batch_equals_one_at_a_time(device):
    from nan_amd.batch import SceneBatch

    _say("eager batch vs single scenes")
    nodes, kws, bounces = _small_batch()
    batch = SceneBatch(device, nodes, kws, bounces)
    enc = device.create_command_encoder()
    for it in range(3):
        batch.execute(enc, it, reset=(it == 0))
    _assert_same(batch, _single_runs(batch, 3), "eager")


def test_batch_with_russian_roulette_equals_one_at_a_time(device):
    from nan_amd.batch import SceneBatch

    _say("eager batch with russian roulette on three of five scenes vs single scenes")
    nodes, kws, bounces = _small_batch()
    for k in (0, 1, 3):
        nodes[k].russian_roulette = True
    bounces = [8, 8, 3, 6, 2]
    batch = SceneBatch(device, nodes, kws, bounces)
    enc = device.create_command_encoder()
    for it in range(2):
        batch.execute(enc, it, reset=(it == 0))
    _assert_same(batch, _single_runs(batch, 2), "roulette")


def test_batch_graph_replays_equal_eager(device):
    from nan_amd.batch import SceneBatch

    _say("graph replays vs single scenes")
    nodes, kws, bounces = _small_batch()
    batch = SceneBatch(device, nodes, kws, bounces)
    enc = device.create_command_encoder()
    batch.execute(enc, 0, reset=True)
    torch.cuda.synchronize()
    graph = batch.capture(enc, 1)
    singles_by_frames = {}
    for frames in (2, 3, 4):
        graph.replay()
        torch.cuda.synchronize()
        # eager work between replays
        scratch = torch.randn(1 << 20, device=device.torch_device).sum().item()
        assert np.isfinite(scratch)
        singles_by_frames[frames] = _single_runs(batch, frames)
        _assert_same(batch, singles_by_frames[frames], f"replay to iteration {frames - 1}")
    batch.clear_device_frame()


@pytest.mark.parametrize("tuning", [(0, 16, 16), (2, 24, 8), (3, 32, 32)])
def test_batch_results_do_not_depend_on_launch_shape(device, tuning):
    from nan_amd.batch import SceneBatch

    _say(f"launch shape {tuning}")
    nodes, kws, bounces = _small_batch()
    batch = SceneBatch(device, nodes, kws, bounces)
    batch.tune(*tuning)
    enc = device.create_command_encoder()
    for it in range(2):
        batch.execute(enc, it, reset=(it == 0))
    _assert_same(batch, _single_runs(batch, 2), f"tuning {tuning}")


def test_autotuner_picks_a_measured_candidate(device):
    from nan_amd.autotune import autotune
    from nan_amd.batch import SceneBatch

    _say("autotune")
    nodes, kws, bounces = _small_batch()
    batch = SceneBatch(device, nodes, kws, bounces)
    res = autotune(batch, device.create_command_encoder(),
                   space={"blocks_per_cu": (0, 2), "lds_stack": (16, 32), "refill_steps": (8, 16)}, frames=1)
    assert len(res["results"]) == 8
    assert res["best"]["ms_per_frame"] == min(r["ms_per_frame"] for r in res["results"])
    assert batch.tuning == (res["best"]["blocks_per_cu"], res["best"]["lds_stack"], res["best"]["refill_steps"])


def test_batch_of_one_scene_equals_the_scene_alone(device):
    from nan_amd.batch import SceneBatch

    _say("batch of one scene")
    batch = SceneBatch(device, [batch_scene(41, 6000)],
                       [dict(texels_per_unit=4.0, min_resolution=1, max_resolution=8, max_total_texels=500_000)], [3])
    enc = device.create_command_encoder()
    for it in range(2):
        batch.execute(enc, it, reset=(it == 0))
    _assert_same(batch, _single_runs(batch, 2), "batch of one")
