"""GPU parity: every kernel of libnan_hip against the CPU oracle, through the
C ABI, on seeded inputs.  Integer/byte/index outputs and the FP32 tracer
payload (which follows the same arithmetic contract as the oracle) must be
bit-exact; the tone map (powf/exp2f) is checked to 2e-6 absolute."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle  # noqa: E402
from nan_amd.mesh_colors import MeshColorsLayout  # noqa: E402
from nan_amd.scene_arrays import flatten_scene  # noqa: E402
from nan_amd.scenes import batch_scene, cornell_box, synthetic_scene  # noqa: E402

# the reference's seven RGB9E5 vectors and the words its float64 definition
# assigns them (tests/golden/make_rgb9e5_vectors.py writes this fixture)
_GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "rgb9e5_vectors.npz")
REFERENCE_VECTORS, REFERENCE_WORDS = _GOLDEN["vectors"], _GOLDEN["words"]


def _say(msg: str) -> None:
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nan(device):
    from nan_amd import passes, scene, texture_space_path_tracer

    return scene, texture_space_path_tracer, passes


def _make(device, node, nan, mode="wavefront", spp=1, **layout_kw):
    scene_mod, tst_mod, _ = nan
    layout = MeshColorsLayout.build(node, **layout_kw)
    sc = scene_mod.Scene(device, node)
    tracer = tst_mod.TextureSpacePathTracer(device, sc, layout, mode=mode, samples_per_texel=spp)
    return layout, sc, tracer


# ------------------------------------------------------------------ RGB9E5
def test_rgb9e5_reference_vectors(device, nan):
    _say("rgb9e5 reference vectors")
    _, _, passes = nan
    payload = np.zeros((len(REFERENCE_VECTORS), 4), dtype=np.float32)
    payload[:, :3] = REFERENCE_VECTORS
    src = torch.from_numpy(payload).to(device.torch_device)
    dst = torch.zeros(len(payload), dtype=torch.int32, device=device.torch_device)
    packer = passes.MeshColorsRGB9E5Packer(device)
    enc = device.create_command_encoder()
    packer.execute(enc, src, dst, len(payload))
    got = dst.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, REFERENCE_WORDS)
    dec = packer.unpack(enc, dst, len(payload)).cpu().numpy()[:, :3]
    for w, d in zip(REFERENCE_WORDS, dec):
        np.testing.assert_array_equal(d, oracle.unpack_rgb9e5(int(w)))


def _oracle_red_range(start: int, count: int, parts: int = 8) -> np.ndarray:
    from concurrent.futures import ThreadPoolExecutor

    step = count // parts
    with ThreadPoolExecutor(parts) as pool:
        chunks = list(pool.map(lambda k: oracle.pack_rgb9e5_red_range(start + k * step, step), range(parts)))
    return np.concatenate(chunks)


def test_rgb9e5_exhaustive_single_channel(device, nan):
    """Every float32 bit pattern whose magnitude can produce a non-trivial word
    (biased exponents 101..143, i.e. 2^-26 .. 2^17, both signs) in the red
    channel, plus 2^22 random patterns over the whole 32-bit space, against the
    oracle, which the CPU suite pins to the float64 definition."""
    _say("rgb9e5 exhaustive red channel")
    _, _, passes = nan
    packer = passes.MeshColorsRGB9E5Packer(device)
    enc = device.create_command_encoder()
    chunk = 1 << 23
    dev = device.torch_device
    payload = torch.zeros((chunk, 4), dtype=torch.float32, device=dev)
    out = torch.empty(chunk, dtype=torch.int32, device=dev)
    base = torch.arange(chunk, dtype=torch.int64, device=dev)
    starts = [e << 23 for e in range(101, 144)] + [(1 << 31) | (e << 23) for e in (101, 120, 127, 143)]
    for start in starts:
        bits = base + start
        bits = torch.where(bits >= 2**31, bits - 2**32, bits).to(torch.int32)
        payload[:, 0] = bits.view(torch.float32)
        packer.execute(enc, payload, out, chunk)
        got = out.cpu().numpy().view(np.uint32)
        np.testing.assert_array_equal(got, _oracle_red_range(start, chunk))
    rnd = np.random.default_rng(8).integers(0, 1 << 32, size=1 << 22, dtype=np.uint64).astype(np.uint32)
    vals = np.zeros((len(rnd), 4), dtype=np.float32)
    vals[:, 0] = rnd.view(np.float32)
    src = torch.from_numpy(vals).to(dev)
    dst = torch.empty(len(rnd), dtype=torch.int32, device=dev)
    packer.execute(enc, src, dst, len(rnd))
    np.testing.assert_array_equal(dst.cpu().numpy().view(np.uint32), oracle.pack_rgb9e5(vals[:, :3]))


def test_rgb9e5_random_triplets(device, nan):
    _say("rgb9e5 random triplets")
    _, _, passes = nan
    rng = np.random.default_rng(3)
    n = 1 << 20
    vals = (rng.standard_normal((n, 3)) * 10.0 ** rng.integers(-12, 6, size=(n, 1))).astype(np.float32)
    payload = np.zeros((n, 4), dtype=np.float32)
    payload[:, :3] = vals
    src = torch.from_numpy(payload).to(device.torch_device)
    dst = torch.zeros(n, dtype=torch.int32, device=device.torch_device)
    passes.MeshColorsRGB9E5Packer(device).execute(device.create_command_encoder(), src, dst, n)
    np.testing.assert_array_equal(dst.cpu().numpy().view(np.uint32), oracle.pack_rgb9e5(vals))


# ------------------------------------------------------------------ ray-gen
@pytest.mark.parametrize("iteration", [0, 7])
def test_texel_raygen_bit_exact(device, nan, iteration):
    _say(f"raygen iteration {iteration}")
    node = cornell_box()
    layout, sc, tracer = _make(device, node, nan)
    out, rng = tracer.raygen(device.create_command_encoder(), iteration)
    ref, ref_rng = oracle.texel_raygen(layout.face_table, layout.instance_table, iteration, 1)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref)
    np.testing.assert_array_equal(rng.cpu().numpy().view(np.uint32), ref_rng)


# ------------------------------------------------------------------ hits
def _rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8))
    r[:, 0:3] = o
    r[:, 4:7] = d
    r[:, 7] = 1000.0
    return r.astype(np.float32)


@pytest.mark.parametrize("which", ["cornell", "batch", "synthetic_small"])
def test_trace_rays_bit_exact_vs_oracle(device, nan, which):
    _say(f"trace_rays {which}")
    rng = np.random.default_rng(9)
    if which == "cornell":
        node = cornell_box()
        rays = _rays(rng, 20000, [-0.95, 0.05, -0.95], [0.95, 1.95, 0.95])
    elif which == "batch":
        node = batch_scene(4, 20000)
        rays = _rays(rng, 20000, [-8, 0, -8], [8, 4, 8])
    else:
        node = synthetic_scene(1, terrain_quads=96, box_grid=12, spheres=30, extent=24.0)
        rays = _rays(rng, 20000, [-12, 0, -12], [12, 6, 12])
    scene_mod, _, _ = nan
    sc = scene_mod.Scene(device, node)
    hits, ids, occ = sc.trace_rays(torch.from_numpy(rays).to(device.torch_device))
    orc = oracle.OracleScene(flatten_scene(node))
    t, b, inst, prim, o = orc.intersect(rays)
    h = hits.cpu().numpy()
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32)[:, 0], inst)
    np.testing.assert_array_equal(ids.cpu().numpy().view(np.uint32)[:, 1], prim)
    np.testing.assert_array_equal(h[:, 0].view(np.uint32), t.view(np.uint32))
    hit = inst != 0xFFFFFFFF
    np.testing.assert_array_equal(h[hit, 1:3].view(np.uint32), b[hit].view(np.uint32))
    np.testing.assert_array_equal(occ.cpu().numpy().astype(np.uint32), o)
    assert hit.mean() > 0.3


# ------------------------------------------------------------------ tracer
def _oracle_payload(node, layout, iterations, bounces=3, reset_first=True, spp=1):
    orc = oracle.OracleScene(flatten_scene(node))
    p = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    counts = np.zeros(4, dtype=np.uint64)
    for it in range(iterations):
        counts += orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, it, spp, bounces,
                                   reset_first and it == 0, threads=oracle.cpu_count())
    return p, counts


@pytest.mark.parametrize("mode", ["wavefront", "megakernel"])
def test_tracer_cornell_bit_exact(device, nan, mode):
    _say(f"tracer cornell 3 iterations ({mode})")
    node = cornell_box()
    layout, sc, tracer = _make(device, node, nan, mode=mode)
    dev = device.torch_device
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
    counters = torch.zeros(8, dtype=torch.int64, device=dev)
    enc = device.create_command_encoder()
    for it in range(3):
        if it == 2:
            tracer.execute(enc, payload, it, reset=False, counters=counters)
        else:
            tracer.execute(enc, payload, it, reset=(it == 0))
    got = payload.cpu().numpy()
    ref, _ = _oracle_payload(node, layout, 3)
    mism = np.nonzero(np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1))[0]
    assert len(mism) == 0, f"{len(mism)} slots differ, first {mism[:10]}"
    # counters of the last iteration equal the oracle's
    orc = oracle.OracleScene(flatten_scene(node))
    p = ref.copy()
    cnt = orc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, 2, 1, 3, False)
    c = counters.cpu().numpy()
    assert c[0] == cnt[0] and c[1] == cnt[1] and c[4] == cnt[3]


@pytest.mark.parametrize("mode", ["wavefront", "megakernel"])
@pytest.mark.parametrize("which", ["batch", "synthetic_small"])
def test_tracer_sunlit_scene_bit_exact(device, nan, which, mode):
    _say(f"tracer {which} ({mode})")
    if which == "batch":
        node = batch_scene(21, 6000)
        kw = dict(texels_per_unit=8.0, min_resolution=1, max_resolution=8, max_total_texels=2_000_000)
    else:
        node = synthetic_scene(2, terrain_quads=64, box_grid=8, spheres=20, extent=20.0)
        kw = dict(texels_per_unit=8.0, min_resolution=2, max_resolution=8, max_total_texels=4_000_000)
    layout, sc, tracer = _make(device, node, nan, mode=mode, **kw)
    dev = device.torch_device
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
    enc = device.create_command_encoder()
    counters = torch.zeros(8, dtype=torch.int64, device=dev)
    for it in range(2):
        tracer.execute(enc, payload, it, reset=(it == 0), counters=counters if it == 1 else None)
    got = payload.cpu().numpy()
    ref, _ = _oracle_payload(node, layout, 1)
    p = ref.copy()
    cnt = oracle.OracleScene(flatten_scene(node)).trace_texels(
        layout.face_table, layout.instance_table, layout.side_table, p, 1, 1, 3, False, threads=oracle.cpu_count())
    diff = np.any(got.view(np.uint32) != p.view(np.uint32), axis=1)
    assert diff.sum() == 0, f"{diff.sum()} of {len(diff)} slots differ"
    c = counters.cpu().numpy()
    assert (c[0], c[1], c[4]) == (cnt[0], cnt[1], cnt[3])


@pytest.mark.parametrize("mode", ["wavefront", "megakernel"])
def test_tracer_russian_roulette_bit_exact(device, nan, mode):
    # roulette on a sunlit scene with paths of up to 6 bounces: the HIP path
    # and the oracle draw the same survival numbers and end the same paths
    _say(f"tracer russian roulette, 6 bounces ({mode})")
    node = batch_scene(23, 5000)
    node.russian_roulette = True
    kw = dict(texels_per_unit=6.0, min_resolution=1, max_resolution=8, max_total_texels=2_000_000)
    scene_mod, tst_mod, _ = nan
    layout = MeshColorsLayout.build(node, **kw)
    sc = scene_mod.Scene(device, node)
    tracer = tst_mod.TextureSpacePathTracer(device, sc, layout, mode=mode, max_bounces=6)
    dev = device.torch_device
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
    counters = torch.zeros(8, dtype=torch.int64, device=dev)
    enc = device.create_command_encoder()
    for it in range(2):
        tracer.execute(enc, payload, it, reset=(it == 0), counters=counters if it == 1 else None)
    got = payload.cpu().numpy()
    ref, _ = _oracle_payload(node, layout, 1, bounces=6)
    cnt = oracle.OracleScene(flatten_scene(node)).trace_texels(
        layout.face_table, layout.instance_table, layout.side_table, ref, 1, 1, 6, False, threads=oracle.cpu_count())
    diff = np.any(got.view(np.uint32) != ref.view(np.uint32), axis=1)
    assert diff.sum() == 0, f"{diff.sum()} of {len(diff)} slots differ"
    c = counters.cpu().numpy()
    assert (c[0], c[1], c[4]) == (cnt[0], cnt[1], cnt[3])
    # roulette ends paths early: fewer extension rays than 6 per path
    assert c[0] < 6 * c[4]


def test_wavefront_two_samples_per_texel_bit_exact(device, nan):
    _say("tracer spp=2 wavefront vs megakernel vs oracle")
    node = batch_scene(22, 4000)
    kw = dict(texels_per_unit=6.0, min_resolution=1, max_resolution=8, max_total_texels=2_000_000)
    outs = []
    for mode in ("wavefront", "megakernel"):
        layout, sc, tracer = _make(device, node, nan, mode=mode, spp=2, **kw)
        payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
        enc = device.create_command_encoder()
        for it in range(2):
            tracer.execute(enc, payload, it, reset=(it == 0))
        outs.append(payload.cpu().numpy())
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    ref, _ = _oracle_payload(node, layout, 2, spp=2)
    np.testing.assert_array_equal(outs[0].view(np.uint32), ref.view(np.uint32))


def test_two_stream_sweeps_equal_one_stream_order(device, nan):
    # the double-sided texels' sweeps run on a second stream unless stage
    # timing is on; both orders leave the same bits and trace the same rays
    # (node visits and triangle tests depend on which rays share a wave, as
    # for any launch shape, so they are not compared)
    _say("wavefront two-stream vs one-stream order")
    node = synthetic_scene(3, terrain_quads=48, box_grid=8, spheres=24, extent=20.0)
    kw = dict(texels_per_unit=8.0, min_resolution=2, max_resolution=8, max_total_texels=4_000_000)
    layout, sc, tracer = _make(device, node, nan, **kw)
    assert layout.total_payload_count > layout.total_surface_texels  # some double-sided instances
    dev = device.torch_device
    outs, counts = [], []
    for profiled in (False, True):
        tracer.profile(profiled)
        payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
        counters = torch.zeros(8, dtype=torch.int64, device=dev)
        enc = device.create_command_encoder()
        for it in range(3):
            tracer.execute(enc, payload, it, reset=(it == 0), counters=counters if it == 2 else None)
        outs.append(payload.cpu().numpy())
        counts.append(counters.cpu().numpy())
    tracer.profile(False)
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    for k in (0, 1, 4):  # extend rays, shadow rays, slot updates
        assert counts[0][k] == counts[1][k], (k, counts[0][k], counts[1][k])


def test_stage_timing_reports_every_stage(device, nan):
    node = cornell_box()
    layout, sc, tracer = _make(device, node, nan)
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    tracer.profile(True)
    tracer.execute(device.create_command_encoder(), payload, 0, reset=True)
    st = tracer.stage_times()
    tracer.profile(False)
    # front + back phases: 2 generate, 6 extend, 6 shade, 6 shadow, 2 finish
    assert {k: v[1] for k, v in st.items()} == {"generate": 2, "extend": 6, "shade": 6, "shadow": 6, "finish": 2}
    assert all(v[0] > 0 for v in st.values())


# ------------------------------------------------------------------ resolve / accumulate / tonemap
def test_resolve_bit_exact_live_and_frozen(device, nan):
    _say("resolve")
    _, _, passes = nan
    node = cornell_box()
    layout, sc, tracer = _make(device, node, nan)
    dev = device.torch_device
    payload = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=dev)
    enc = device.create_command_encoder()
    tracer.execute(enc, payload, 0, reset=True)
    cam = node.camera
    cam.begin_frame(128, 96)
    out = torch.zeros((96, 128, 4), dtype=torch.float32, device=dev)
    resolve = passes.MeshColorsResolve(device, sc, layout, tracer)
    resolve.execute(enc, payload, out)
    orc = oracle.OracleScene(flatten_scene(node))
    p = payload.cpu().numpy()
    ref = orc.resolve(layout.face_table, layout.instance_table, layout.side_table, cam.pack(), 128, 96, payload=p)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    packed = torch.zeros(layout.total_payload_count, dtype=torch.int32, device=dev)
    passes.MeshColorsRGB9E5Packer(device).execute(enc, payload, packed, layout.total_payload_count)
    resolve.execute_frozen(enc, packed, out)
    ref = orc.resolve(layout.face_table, layout.instance_table, layout.side_table, cam.pack(), 128, 96,
                      packed=packed.cpu().numpy().view(np.uint32))
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_accumulate_and_tonemap(device, nan):
    _say("accumulate + tonemap")
    _, _, passes = nan
    from nan_amd.render_data import RenderData

    dev = device.torch_device
    rng = np.random.default_rng(4)
    rd = RenderData(device)
    acc = passes.Accumulator(device)
    tm = passes.ToneMapper(device, exposure=0.5)
    enc = device.create_command_encoder()
    accum_ref = np.zeros((32, 48, 4), dtype=np.float32)
    out = torch.zeros((32, 48, 4), dtype=torch.float32, device=dev)
    for k in range(3):
        img = (rng.random((32, 48, 4)) * 20).astype(np.float32)
        acc.execute(enc, rd, torch.from_numpy(img).to(dev), out, reset=(k == 0))
        ref = oracle.accumulate(img, accum_ref, reset=(k == 0))
        np.testing.assert_array_equal(out.cpu().numpy(), ref)
    tmo = torch.zeros_like(out)
    tm.execute(enc, out, tmo)
    np.testing.assert_allclose(tmo.cpu().numpy(), oracle.tonemap(out.cpu().numpy(), 0.5), atol=2e-6)


def test_renderer_frames_and_freeze(device, nan):
    _say("renderer frames")
    scene_mod, _, _ = nan
    from nan_amd.render_data import RenderData
    from nan_amd.texture_space_path_tracing_renderer import TextureSpacePathTracingRenderer

    node = cornell_box()
    sc = scene_mod.Scene(device, node)
    r = TextureSpacePathTracingRenderer()
    r.initialize(device, sc)
    rd = RenderData(device)
    out = torch.zeros((64, 64, 4), dtype=torch.float32, device=device.torch_device)
    for frame in range(4):
        node.camera.begin_frame(64, 64)
        r.render(device.create_command_encoder(), out, frame, device, sc, rd)
    assert r.texture_iteration == 4
    img = out.cpu().numpy()
    assert np.isfinite(img).all() and img[..., :3].max() > 0.05
    # the save runs 5 surface-filter passes, then packs (texture_space_path_tracing_renderer.py:181-231)
    from nan_amd.mesh_colors_surface_filter import face_normals
    from nan_amd.texture_space_path_tracing_renderer import MESH_COLORS_BUFFER_KEY
    from oracle import filter_oracle

    cache = rd.get_buffer(MESH_COLORS_BUFFER_KEY, struct_size=16,
                          element_count=r.layout.total_payload_count).cpu().numpy().view(np.float32).reshape(-1, 4)
    r._request_rgb9e5_save()
    r.render(device.create_command_encoder(), out, 5, device, sc, rd)
    assert r._frozen_rgb9e5 is not None and r.texture_iteration == 4
    lay = r.layout
    normals = face_normals(node, lay)
    want = cache
    for _ in range(5):
        want = filter_oracle.surface_filter(lay.face_table, lay.instance_table, lay.side_table, lay.adjacency_table,
                                            normals, want, np.exp(-0.5), np.radians(30.0))
    words = r._frozen_rgb9e5.cpu().numpy().view(np.uint32).reshape(-1)[:lay.total_payload_count]
    got = np.stack([oracle.unpack_rgb9e5(int(w))[:3] for w in words[::97]])
    ref = want[::97, :3].astype(np.float64)
    # RGB9E5 keeps 9 mantissa bits of the largest channel
    tol = np.maximum(ref.max(axis=1, keepdims=True), 1e-6) * 2.0 ** -8
    assert (np.abs(got - ref) <= tol).all()
    # write the frozen cache to disk, load it into a fresh renderer: same frozen resolve
    import tempfile
    from pathlib import Path

    node.camera.begin_frame(64, 64)
    r.reset_accumulator = True  # one frame of the frozen resolve, same camera jitter for both renderers
    r.render(device.create_command_encoder(), out, 6, device, sc, rd)
    frozen_img = out.clone()
    with tempfile.TemporaryDirectory() as tmp:
        path = Path(tmp) / "cornell.nmc"
        r.write_frozen_cache(path)
        r2 = TextureSpacePathTracingRenderer()
        r2.initialize(device, sc)
        rd2 = RenderData(device)
        r2.read_frozen_cache(path, rd2)
    out2 = torch.zeros_like(out)
    r2.render(device.create_command_encoder(), out2, 6, device, sc, rd2)
    assert r2.texture_iteration == 0 and r2._frozen_rgb9e5 is not None
    assert torch.equal(out2, frozen_img)
