"""An imported scene on the GPU: the reference's demo scene written as GLB
and read back by ``SceneNode.load_asset`` traces, in both tracer modes, to
the same payload bits as the procedural scene and as the CPU oracle."""

from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
import asset_files as W  # noqa: E402


def _say(msg: str) -> None:
    print(f"[gpu-asset-import] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", ["wavefront", "megakernel"])
def test_imported_demo_traces_like_the_procedural_scene(device, tmp_path, mode):
    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.scene import Scene
    from nan_amd.scene_arrays import flatten_scene
    from nan_amd.scene_node import SceneNode
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say(f"demo GLB round trip, {mode}")
    demo = SceneNode.demo()
    doc, blob = W.gltf_document(demo)
    W.write_glb(str(tmp_path / "demo.glb"), doc, blob)
    imported = SceneNode.load_asset(str(tmp_path / "demo.glb"))
    got = []
    for node in (demo, imported):
        layout = MeshColorsLayout.build(node, texels_per_unit=16.0, min_resolution=4, max_resolution=64)
        sc = Scene(device, node)
        tracer = TextureSpacePathTracer(device, sc, layout, mode=mode)
        payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
        enc = device.create_command_encoder()
        for it in range(2):
            tracer.execute(enc, payload, it, reset=(it == 0))
        got.append(payload.cpu().numpy())
    np.testing.assert_array_equal(got[0].view(np.uint32), got[1].view(np.uint32))
    layout = MeshColorsLayout.build(imported, texels_per_unit=16.0, min_resolution=4, max_resolution=64)
    ref = np.zeros_like(got[1])
    orc = oracle.OracleScene(flatten_scene(imported))
    for it in range(2):
        orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 1, 3, it == 0,
                         threads=oracle.cpu_count())
    np.testing.assert_array_equal(got[1].view(np.uint32), ref.view(np.uint32))
    assert got[1][:, :3].max() > 0.0
