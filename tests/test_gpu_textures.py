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


def test_imported_textured_glb_bit_exact(device, tmp_path):
    """The reference's demo scene written as a GLB with embedded PNG
    textures (base colour with alpha, normal map, emission) and mask / blend
    materials, imported and traced: wavefront payload == oracle."""
    import os

    sys.path.insert(0, os.path.dirname(__file__))
    import asset_files as W
    from nan_amd.scene import Scene
    from nan_amd.scene_node import SceneNode
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    rng = np.random.default_rng(4)
    demo = SceneNode.demo()
    doc, blob = W.gltf_document(demo)
    base = rng.integers(0, 256, size=(16, 12, 4), dtype=np.uint8)
    normal = rng.integers(60, 256, size=(8, 8, 4), dtype=np.uint8)
    emit = rng.integers(0, 256, size=(4, 4, 4), dtype=np.uint8)
    blob = W.add_textures(doc, blob, {(0, "baseColorTexture"): base, (0, "normalTexture"): normal,
                                      (1, "emissiveTexture"): emit, (1, "baseColorTexture"): base})
    doc["materials"][0]["alphaMode"] = "MASK"
    doc["materials"][1]["alphaMode"] = "BLEND"
    doc["materials"][1]["emissiveFactor"] = [2.0, 1.0, 0.5]
    W.write_glb(str(tmp_path / "demo_tex.glb"), doc, blob)
    node = SceneNode.load_asset(str(tmp_path / "demo_tex.glb"))
    _say("imported textured GLB: wavefront vs oracle")
    layout = MeshColorsLayout.build(node, texels_per_unit=8.0, min_resolution=2, max_resolution=8)
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3)
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    enc = device.create_command_encoder()
    orc = oracle.OracleScene(flatten_scene(node))
    ref = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    for it in range(2):
        sc.update()
        tracer.execute(enc, payload, it, reset=(it == 0))
        orc.set_frame_index(sc.frame_index)
        orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 1, 3, it == 0,
                         threads=oracle.cpu_count())
    _bits_equal(payload.cpu().numpy(), ref, "imported GLB")


@pytest.mark.parametrize("texel_query", [False, True])
def test_learned_forms_on_textured_scene_wavefront_equals_per_texel(device, texel_query):
    """The learned forms' material-feature kernels: on the textured scene the
    wavefront queues and the per-texel kernels write identical training
    records and (fixed random cache) identical payloads."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say(f"learned on textured scene, texel_query={texel_query}")
    node = textured_scene(3)
    layout = MeshColorsLayout.build(node, **LAYOUT)
    sc = Scene(device, node)
    sc.set_frame_index(2)
    tracers = {m: TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode=m)
               for m in ("megakernel", "wavefront")}
    n = min(40_000, layout.total_surface_texels)
    recs = {m: t.learned_training_records(5, n, texel_query=texel_query) for m, t in tracers.items()}
    for k in ("pos", "nrm", "alb", "target", "valid"):
        a, b = recs["megakernel"][k].cpu().numpy(), recs["wavefront"][k].cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    cache = RadianceCache.for_scene(device, sc)
    w, t = cache.get_params()
    rng = np.random.default_rng(12)
    cache.set_params((rng.standard_normal(w.shape) * 0.1).astype(w.dtype),
                     (rng.standard_normal(t.shape) * 0.1).astype(t.dtype))
    outs = {}
    for m, tr in tracers.items():
        out = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
        enc = device.create_command_encoder()
        for it in range(2):
            tr.execute_learned(enc, out, it, cache, reset=(it == 0), train_samples=0, texel_query=texel_query)
        outs[m] = out.cpu().numpy()
    _bits_equal(outs["wavefront"], outs["megakernel"], "learned payload")
    assert np.isfinite(outs["wavefront"]).all() and outs["wavefront"][:, :3].max() > 0.0


def _lamp_scene():
    """A textured floor under a 1x1 lamp facing down whose emission is a
    checker texture times (4, 4, 4) and whose blend alpha cuts about half of
    it away; no sun and no sky, so the lamp is the only light."""
    from nan_amd.material import ALPHA_MODE_BLEND, Material
    from nan_amd.mesh import Mesh
    from nan_amd.scene_node import Lighting, SceneNode
    from nan_amd.transform import Transform

    node = SceneNode()
    node.lighting = Lighting.black()
    floor_tex = np.zeros((4, 4, 4), dtype=np.uint8)
    floor_tex[...] = (180, 150, 90, 255)
    floor = node.add_material(Material(base_color_texture=floor_tex))
    glow = np.zeros((8, 8, 4), dtype=np.uint8)
    glow[..., :3] = np.where((np.add.outer(np.arange(8), np.arange(8)) % 2 == 0)[..., None], 250, 40)
    glow[..., 3] = 255
    cut = np.full((8, 8, 4), 255, dtype=np.uint8)
    cut[..., 3] = 128
    lamp = node.add_material(Material(base_color_texture=cut, emissive=(4.0, 4.0, 4.0), emissive_texture=glow,
                                      alpha_mode=ALPHA_MODE_BLEND))
    node.add_instance(node.add_mesh(Mesh.create_quad((4.0, 4.0))), floor, node.add_transform(Transform()))
    node.add_instance(node.add_mesh(Mesh.create_quad((1.0, 1.0))), lamp,
                      node.add_transform(Transform((0.0, 1.0, 0.0), (180.0, 0.0, 0.0), (1.0, 1.0, 1.0))))
    return node


def test_learned_light_sampling_sees_emissive_texture_and_alpha(device):
    """Light sampling of a textured, blend-alpha emitter (its emission at the
    sampled point, zero where the alpha test cuts it away) is unbiased: with
    a zero cache and one bounce, the floor's mean irradiance with emitter
    light sampling equals the one without (cosine-sampled hits only) to the
    noise of 256 frames (the texture averages 0.49 of the factor and the
    cut keeps about half, so missing either would show)."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    node = _lamp_scene()
    layout = MeshColorsLayout.build(node, texels_per_unit=8.0, min_resolution=2, max_resolution=16)
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=1, mode="wavefront")
    cache = RadianceCache.for_scene(device, sc)
    w, t = cache.get_params()
    cache.set_params(np.zeros_like(w), np.zeros_like(t))
    enc = device.create_command_encoder()
    means = {}
    for nee in (True, False):
        _say(f"lamp scene, emitter light sampling {nee}")
        out = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
        for it in range(256):
            sc.set_frame_index(it)
            tracer.execute_learned(enc, out, it, cache, reset=(it == 0), train_samples=0, emitter_nee=nee,
                                   ramped_mean=False)
        e = out[:, :3].cpu().numpy().astype(np.float64)
        assert np.isfinite(e).all()
        means[nee] = e.mean()
    _say(f"means: light sampling {means[True]:.5f}, bounce only {means[False]:.5f}")
    assert means[False] > 0.05
    assert abs(means[True] - means[False]) < 0.03 * means[False]
