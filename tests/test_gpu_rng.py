"""The production random streams (rng_mode "pcg") on the GPU: payloads bit
for bit equal to the oracle's in both tracer modes, with several samples per
texel, in batches and on textured scenes; the learned forms' records equal
between the wavefront and the per-texel kernels; and the PCG streams
converge to a reference-stream bake as fast as the reference's own stream.
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
from nan_amd.scenes import batch_scene, textured_scene  # noqa: E402

LAYOUT = dict(texels_per_unit=4.0, min_resolution=1, max_resolution=4, max_total_texels=400_000)


def _say(msg: str) -> None:
    print(f"[gpu-rng] {msg}", flush=True)
    sys.stdout.flush()


def _bits_equal(got: np.ndarray, ref: np.ndarray, tag: str) -> None:
    g, r = got.view(np.uint32), ref.view(np.uint32)
    bad = np.nonzero(np.any(g != r, axis=1))[0]
    assert len(bad) == 0, f"{tag}: {len(bad)} of {len(g)} slots differ, first {bad[0]}"


@pytest.mark.parametrize("mode", ["wavefront", "megakernel"])
@pytest.mark.parametrize("which", ["batch", "textured"])
def test_pcg_payload_bit_exact(device, mode, which):
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    node = batch_scene(21, 3000) if which == "batch" else textured_scene(3)
    layout = MeshColorsLayout.build(node, **LAYOUT)
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, mode=mode, samples_per_texel=2, max_bounces=3, rng="pcg")
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    enc = device.create_command_encoder()
    orc = oracle.OracleScene(flatten_scene(node))
    orc.set_rng("pcg")
    ref = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    for it in range(2):
        _say(f"pcg {which} {mode} iteration {it}")
        tracer.execute(enc, payload, it, reset=(it == 0))
        orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 2, 3, it == 0,
                         threads=oracle.cpu_count())
    _bits_equal(payload.cpu().numpy(), ref, f"pcg {which} {mode}")
    # a different stream than the reference's
    tracer.set_rng("reference")
    other = torch.zeros_like(payload)
    tracer.execute(enc, other, 0, reset=True)
    orc.set_rng("reference")
    ref0 = np.zeros_like(ref)
    orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref0, 0, 2, 3, True,
                     threads=oracle.cpu_count())
    _bits_equal(other.cpu().numpy(), ref0, f"reference {which} {mode}")


def test_pcg_batch_bit_exact(device):
    from nan_amd.batch import SceneBatch

    nodes = [batch_scene(21, 3000), textured_scene(4, 3000), batch_scene(22, 1500)]
    _say("pcg batch of 3 scenes")
    batch = SceneBatch(device, nodes, [LAYOUT] * 3, [3, 2, 4], rng="pcg")
    enc = device.create_command_encoder()
    for it in range(2):
        batch.execute(enc, it, reset=(it == 0))
    torch.cuda.synchronize()
    for k, (node, layout) in enumerate(zip(nodes, batch.layouts)):
        orc = oracle.OracleScene(flatten_scene(node))
        orc.set_rng("pcg")
        ref = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        for it in range(2):
            orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 1,
                             batch.max_bounces[k], it == 0, threads=oracle.cpu_count())
        _bits_equal(batch.payloads[k].cpu().numpy(), ref, f"pcg batch scene {k}")


def test_pcg_learned_records_wavefront_equal_per_texel(device):
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    node = batch_scene(21, 3000)
    layout = MeshColorsLayout.build(node, **LAYOUT)
    sc = Scene(device, node)
    recs = {}
    for m in ("megakernel", "wavefront"):
        tr = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode=m, rng="pcg")
        recs[m] = tr.learned_training_records(3, 20000)
    for k in ("pos", "nrm", "alb", "target", "valid"):
        a, b = recs["megakernel"][k].cpu().numpy(), recs["wavefront"][k].cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    ref = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="wavefront")
    r = ref.learned_training_records(3, 20000)
    assert not np.array_equal(r["target"].cpu().numpy(), recs["wavefront"]["target"].cpu().numpy())


def test_pcg_converges_like_the_reference_stream(device):
    """RMSE after 32 frames against a 1024-frame reference-stream bake: the
    PCG streams are within 10% of the reference stream (same expectation,
    same per-texel variance)."""
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    node = batch_scene(21, 3000)
    layout = MeshColorsLayout.build(node, **LAYOUT)
    sc = Scene(device, node)
    dev = device.torch_device
    enc = device.create_command_encoder()
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3)
    bake = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
    _say("bake, 1024 frames")
    for it in range(1024):
        tracer.execute(enc, bake, 50_000 + it, reset=(it == 0))
    b = bake[:, :3].double()
    rmse = {}
    for rng in ("reference", "pcg"):
        tracer.set_rng(rng)
        out = torch.zeros_like(bake)
        for it in range(32):
            tracer.execute(enc, out, it, reset=(it == 0))
        rmse[rng] = float(torch.sqrt(((out[:, :3].double() - b) ** 2).mean()))
    _say(f"rmse after 32 frames: {rmse}")
    assert abs(rmse["pcg"] / rmse["reference"] - 1.0) < 0.1, rmse
