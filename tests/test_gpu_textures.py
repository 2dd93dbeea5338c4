"""Textured and alpha-tested materials on the GPU against the oracle, bit for
bit: base-colour, emissive and normal textures (bilinear, wrap, sRGB
decode), alpha mask and stochastic blend on every candidate hit
(scene.slang:124-149, 313-368, 1017-1172), through the C ABI.

* texel payloads of the wavefront and megakernel tracers over two
  iterations whose blend frame index changes (Scene.update);
* closest hit and occlusion of random rays (Scene.intersect /
  get_visibility);
* the camera resolve (base-colour and emissive textures at the pixel hit);
* a batch of textured and untextured scenes, each slot equal to the oracle
  of its own scene at the batch's frame index.
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
    print(f"[gpu-textures] {msg}", flush=True)
    sys.stdout.flush()


def _bits_equal(got: np.ndarray, ref: np.ndarray, tag: str) -> None:
    g = got.view(np.uint32).reshape(len(got), -1)
    r = ref.view(np.uint32).reshape(len(ref), -1)
    bad = np.nonzero(np.any(g != r, axis=1))[0]
    assert len(bad) == 0, f"{tag}: {len(bad)} of {len(g)} rows differ, first {bad[0]}: {got[bad[0]]} vs {ref[bad[0]]}"


@pytest.mark.parametrize("mode", ["wavefront", "megakernel"])
def test_textured_scene_payload_bit_exact(device, mode):
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    node = textured_scene(3)
    layout = MeshColorsLayout.build(node, **LAYOUT)
    sc = Scene(device, node)
    assert sc.info.triangle_count > 1000
    tracer = TextureSpacePathTracer(device, sc, layout, mode=mode, samples_per_texel=1, max_bounces=3)
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    enc = device.create_command_encoder()
    orc = oracle.OracleScene(flatten_scene(node))
    ref = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    for it in range(2):
        sc.update()  # frame index 1, then 2
        _say(f"{mode}: iteration {it}, frame index {sc.frame_index}")
        tracer.execute(enc, payload, it, reset=(it == 0))
        orc.set_frame_index(sc.frame_index)
        orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 1, 3, it == 0,
                         threads=oracle.cpu_count())
    got = payload.cpu().numpy()
    _bits_equal(got, ref, mode)
    assert np.isfinite(got).all() and got[:, :3].max() > 0


def _rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8))
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = 1000.0
    return r.astype(np.float32)


def test_textured_trace_rays_bit_exact(device):
    from nan_amd.scene import Scene

    node = textured_scene(5)
    sc = Scene(device, node)
    orc = oracle.OracleScene(flatten_scene(node))
    rng = np.random.default_rng(11)
    rays = _rays(rng, 40000, [-8, 0, -8], [8, 4, 8])
    for frame in (0, 7):
        _say(f"trace_rays, frame index {frame}")
        sc.set_frame_index(frame)
        orc.set_frame_index(frame)
        hits, ids, occ = sc.trace_rays(torch.from_numpy(rays).to(device.torch_device))
        t, b, inst, prim, o = orc.intersect(rays)
        h = hits.cpu().numpy()
        np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32)[:, 0], inst)
        np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32)[:, 1], prim)
        np.testing.assert_array_equal(h[:, 0].view(np.uint32), t.view(np.uint32))
        hit = inst != 0xFFFFFFFF
        np.testing.assert_array_equal(h[hit, 1:3].view(np.uint32), b[hit].view(np.uint32))
        np.testing.assert_array_equal(occ.cpu().numpy().astype(np.uint32), o)
        assert hit.mean() > 0.3


def test_textured_resolve_bit_exact(device):
    from nan_amd import passes
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say("resolve")
    node = textured_scene(3)
    layout = MeshColorsLayout.build(node, **LAYOUT)
    sc = Scene(device, node)
    sc.set_frame_index(4)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3)
    dev = device.torch_device
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
    enc = device.create_command_encoder()
    tracer.execute(enc, payload, 0, reset=True)
    cam = node.camera
    cam.begin_frame(160, 120)
    out = torch.zeros((120, 160, 4), dtype=torch.float32, device=dev)
    passes.MeshColorsResolve(device, sc, layout, tracer).execute(enc, payload, out)
    orc = oracle.OracleScene(flatten_scene(node))
    orc.set_frame_index(4)
    ref = orc.resolve(layout.face_table, layout.instance_table, layout.side_table, cam.pack(), 160, 120,
                      payload=payload.cpu().numpy())
    _bits_equal(out.cpu().numpy().reshape(-1, 4), ref.reshape(-1, 4), "resolve")


def test_textured_batch_bit_exact(device):
    from nan_amd.batch import SceneBatch

    nodes = [textured_scene(3), batch_scene(21, 3000), textured_scene(4, 3000)]
    kws = [LAYOUT] * 3
    _say("batch of 3 scenes (2 textured), frame index 9")
    batch = SceneBatch(device, nodes, kws, [3, 2, 3])
    enc = device.create_command_encoder()
    batch.set_frame_index(enc, 9)
    for it in range(2):
        batch.execute(enc, it, reset=(it == 0))
    torch.cuda.synchronize()
    for k, (node, layout) in enumerate(zip(nodes, batch.layouts)):
        orc = oracle.OracleScene(flatten_scene(node))
        orc.set_frame_index(9)
        ref = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        for it in range(2):
            orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 1,
                             batch.max_bounces[k], it == 0, threads=oracle.cpu_count())
        _bits_equal(batch.payloads[k].cpu().numpy(), ref, f"batch scene {k}")
