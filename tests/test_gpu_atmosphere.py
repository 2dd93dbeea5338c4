"""Atmosphere tables and the physical sky on the GPU against the CPU oracle
(oracle/atmosphere_oracle.c, oracle/nan_oracle.c).

Table generation calls exp/sin/cos/acos/pow, whose GPU and C-library
results differ in the last bits, so each table is compared stage by stage
(every stage fed the same input tables) to a relative tolerance of 1e-4.
Everything downstream of a table (the view-point transmittance, the sky
lookup, the tracer and the resolve) uses only correctly rounded operations
and is compared bit for bit."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from oracle import atmosphere as A  # noqa: E402
from nan_amd.scene_node import Lighting  # noqa: E402
from nan_amd.scenes import synthetic_scene  # noqa: E402

RTOL = 1e-4


def _say(msg: str) -> None:
    print(f"[gpu-atmosphere] {msg}", flush=True)
    sys.stdout.flush()


def _close(got: np.ndarray, want: np.ndarray, what: str) -> None:
    err = np.abs(got.astype(np.float64) - want)
    bound = RTOL * np.abs(want.astype(np.float64)) + 1e-7 * float(np.abs(want).max())
    worst = float((err / np.maximum(bound, 1e-30)).max())
    _say(f"{what}: max |diff| {err.max():.3e}, worst diff/bound {worst:.3f}")
    assert (err <= bound).all(), f"{what}: {(err > bound).sum()} values beyond the tolerance"


@pytest.fixture(scope="module")
def oracle_tables():
    t = A.transmittance_lut()
    m = A.multiscattering_lut(t)
    return t, m


def test_tables_match_oracle_stage_by_stage(device, oracle_tables):
    from nan_amd.atmosphere import AtmosphereMultiScatteringLUT, AtmosphereSkyViewLUT, AtmosphereTransmittanceLUT

    _say("transmittance / multiple scattering / sky view tables")
    t_ref, m_ref = oracle_tables
    dev = device.torch_device
    enc = device.create_command_encoder()
    tg = AtmosphereTransmittanceLUT(device)
    tg.generate(enc)
    _close(tg.get_texture().cpu().numpy(), t_ref, "transmittance")
    mg = AtmosphereMultiScatteringLUT(device)
    mg.generate(enc, torch.from_numpy(t_ref).to(dev))
    _close(mg.get_texture().cpu().numpy(), m_ref, "multiple scattering")
    for sun in ((0.0, 0.5, 0.866), (0.4, 0.1, -0.9), (0.0, -0.05, 1.0)):
        sg = AtmosphereSkyViewLUT(device)
        sg.generate(enc, torch.from_numpy(t_ref).to(dev), torch.from_numpy(m_ref).to(dev), sun)
        _close(sg.get_texture().cpu().numpy(), A.sky_view_lut(t_ref, m_ref, sun), f"sky view {sun}")


def test_view_transmittance_and_lookup_bit_exact(device, oracle_tables):
    from nan_amd.atmosphere import SkyAtmosphere

    _say("view transmittance and sky lookup")
    t_ref, m_ref = oracle_tables
    lighting = Lighting(sun_direction=(0.3, 0.45, 0.8), sun_intensity=2.5, sun_color=(1.0, 0.9, 0.7),
                        atmosphere=True)
    rec = lighting.pack()
    atm = SkyAtmosphere(device, rec[0:3])
    t_gpu = atm.transmittance_lut_gen.get_texture().cpu().numpy()
    got = np.asarray(atm.get_sun_transmittance_at_view(), dtype=np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), A.view_transmittance(t_gpu, rec[0:3]).view(np.uint32))
    sky = atm.sky_view_lut.cpu().numpy()
    rng = np.random.default_rng(3)
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    edges = np.array([[0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 0, 0],
                      [1e-8, 1, 0], [-1e-7, 0, -1], [rec[0], rec[1], rec[2]], [-rec[0], 0, -rec[2]]])
    dirs = np.zeros((len(d) + len(edges), 4), dtype=np.float32)
    dirs[:len(d), :3] = d
    dirs[len(d):, :3] = edges
    out = atm.sample_sky_lut(lighting, torch.from_numpy(dirs).to(device.torch_device)).cpu().numpy()
    want = A.sky_view_lookup(sky, rec, dirs)
    np.testing.assert_array_equal(out.view(np.uint32), want.view(np.uint32))
    assert np.isfinite(out).all() and out[:, :3].max() > 0.0


def _sky_scene():
    node = synthetic_scene(2, terrain_quads=32, box_grid=4, spheres=10, extent=20.0)
    node.lighting = Lighting.reference_default()
    node.camera.position = np.array([0.0, 3.0, 14.0])
    node.camera.target = np.array([0.0, 4.0, 0.0])
    node.camera.recompute()
    return node


@pytest.mark.parametrize("mode", ["wavefront", "megakernel"])
def test_tracer_with_physical_sky_bit_exact(device, mode):
    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say(f"tracer with the sky-view table ({mode})")
    node = _sky_scene()
    layout = MeshColorsLayout.build(node, texels_per_unit=4.0, min_resolution=1, max_resolution=8,
                                    max_total_texels=2_000_000)
    sc = Scene(device, node)
    assert sc.atmosphere is not None
    tracer = TextureSpacePathTracer(device, sc, layout, mode=mode)
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    enc = device.create_command_encoder()
    for it in range(2):
        tracer.execute(enc, payload, it, reset=(it == 0))
    got = payload.cpu().numpy()
    orc = oracle.OracleScene(sc.arrays)
    orc.set_sky_view_lut(sc.atmosphere.sky_view_lut.cpu().numpy())
    ref = np.zeros_like(got)
    first = None
    for it in range(2):
        orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, it, 1, 3, it == 0,
                         threads=oracle.cpu_count())
        first = ref.copy() if first is None else first
    diff = np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1)
    assert diff.sum() == 0, f"{diff.sum()} of {len(diff)} slots differ"
    # the table changes the answer: without it (constant sky of zero) many slots differ
    orc2 = oracle.OracleScene(sc.arrays)
    plain = np.zeros_like(got)
    orc2.trace_texels(layout.face_table, layout.instance_table, layout.side_table, plain, 0, 1, 3, True,
                      threads=oracle.cpu_count())
    assert np.any(plain != first, axis=1).mean() > 0.2


def test_resolve_sees_sky_and_sun_disk_bit_exact(device):
    from nan_amd import passes
    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say("resolve with the physical sky")
    node = _sky_scene()
    layout = MeshColorsLayout.build(node, texels_per_unit=4.0, min_resolution=1, max_resolution=8,
                                    max_total_texels=2_000_000)
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout)
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    enc = device.create_command_encoder()
    tracer.execute(enc, payload, 0, reset=True)
    cam = node.camera
    cam.begin_frame(96, 64)
    out = torch.zeros((64, 96, 4), dtype=torch.float32, device=device.torch_device)
    passes.MeshColorsResolve(device, sc, layout, tracer).execute(enc, payload, out)
    orc = oracle.OracleScene(sc.arrays)
    orc.set_sky_view_lut(sc.atmosphere.sky_view_lut.cpu().numpy())
    ref = orc.resolve(layout.face_table, layout.instance_table, layout.side_table, cam.pack(), 96, 64,
                      payload=payload.cpu().numpy())
    img = out.cpu().numpy()
    np.testing.assert_array_equal(img.view(np.uint32), ref.view(np.uint32))
    # the camera looks slightly up: a good part of the frame is blue sky
    assert np.isfinite(img).all()
    rgb = img[..., :3].reshape(-1, 3)
    assert ((rgb[:, 2] > rgb[:, 0]) & (rgb[:, 2] > 0.05)).mean() > 0.1


def test_switching_the_sky_off_restores_the_constant_sky(device):
    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say("atmosphere on -> off")
    node = _sky_scene()
    layout = MeshColorsLayout.build(node, texels_per_unit=2.0, min_resolution=1, max_resolution=4,
                                    max_total_texels=2_000_000)
    sc = Scene(device, node)
    plain = Lighting.sun_and_sky(30.0, 0.0, sky=0.3)
    sc.set_lighting(plain)
    tracer = TextureSpacePathTracer(device, sc, layout)
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    tracer.execute(device.create_command_encoder(), payload, 0, reset=True)
    orc = oracle.OracleScene(sc.arrays)
    ref = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, ref, 0, 1, 3, True,
                     threads=oracle.cpu_count())
    np.testing.assert_array_equal(payload.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    assert math.isclose(float(sc.arrays["lighting"][12]), 0.3, rel_tol=1e-6)
