"""Pins the CPU oracle: RNG, sin/cos, RGB9E5, lattice lookup, ray/triangle
hits against FP64 brute force, and closed-form lighting invariants."""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from nan_amd.material import Material
from nan_amd.mesh import Mesh
from nan_amd.mesh_colors import MeshColorsLayout, colors_per_patch
from nan_amd.scene_arrays import flatten_scene
from nan_amd.scene_node import Lighting, SceneNode
from nan_amd.scenes import batch_scene, cornell_box
from nan_amd.transform import Transform

GOLDEN = Path(__file__).resolve().parent / "golden"


def _rgb9e5_float64():
    """The float64 RGB9E5 definition restated from the reference's test
    (tests/golden/make_rgb9e5_vectors.py; CPU machines only)."""
    return pytest.importorskip("tests.golden.make_rgb9e5_vectors")


def pack_rgb9e5_reference(v):
    return _rgb9e5_float64().pack_rgb9e5_reference(v)


def pack_rgb9e5_reference_array(v):
    return _rgb9e5_float64().pack_rgb9e5_reference_array(v)


def unpack_rgb9e5_reference(w):
    return _rgb9e5_float64().unpack_rgb9e5_reference(w)


# ---------------------------------------------------------------- RNG
def test_rng_matches_lcg_closed_form():
    # common.slang:164-210: seed x*100591 ^ y*102191 ^ s*103643; A=1664525, C=1013904223
    for x, y, s in [(0, 0, 0), (12345, 7, 3), (2**31, 4096, 77)]:
        vals, state = oracle.rng_stream(x, y, s, 40)
        st = ((x * 100591) ^ (y * 102191) ^ (s * 103643)) & 0xFFFFFFFF
        for v in vals:
            st = (1664525 * st + 1013904223) & 0xFFFFFFFF
            assert v == np.float32((st >> 8) * 2.0 ** -24)
        assert state == st


# ---------------------------------------------------------------- sin/cos
def test_sincos_accuracy():
    u = (np.arange(1 << 16, dtype=np.float64) * 2.0 ** -16).astype(np.float32)
    u = np.concatenate([u, np.random.default_rng(0).random(100_000).astype(np.float32), [1.0 - 2**-24, 0.125, 0.375]])
    u = u.astype(np.float32)
    s, c = oracle.sincos_2pi(u)
    ref_s = np.sin(2 * np.pi * u.astype(np.float64))
    ref_c = np.cos(2 * np.pi * u.astype(np.float64))
    assert np.max(np.abs(s - ref_s)) < 3e-7
    assert np.max(np.abs(c - ref_c)) < 3e-7
    # exact special values
    s, c = oracle.sincos_2pi(np.asarray([0.0, 0.25, 0.5, 0.75], dtype=np.float32))
    np.testing.assert_array_equal(s, [0.0, 1.0, 0.0, -1.0])
    np.testing.assert_array_equal(c, [1.0, 0.0, -1.0, 0.0])


# ---------------------------------------------------------------- RGB9E5
def test_rgb9e5_reference_vectors():
    fixture = np.load(GOLDEN / "rgb9e5_vectors.npz")
    vectors, words = fixture["vectors"], fixture["words"]
    np.testing.assert_array_equal(vectors.view(np.uint32), _rgb9e5_float64().REFERENCE_VECTORS.view(np.uint32))
    got = oracle.pack_rgb9e5(vectors)
    np.testing.assert_array_equal(got, words)
    for v, w in zip(vectors, words):
        assert pack_rgb9e5_reference(v) == w
        np.testing.assert_array_equal(oracle.unpack_rgb9e5(int(w)), unpack_rgb9e5_reference(int(w)))


def test_rgb9e5_vectorised_reference_agrees_with_scalar():
    rng = np.random.default_rng(1)
    vals = (rng.standard_normal((2000, 3)) * 10.0 ** rng.integers(-9, 5, size=(2000, 1))).astype(np.float32)
    vec = pack_rgb9e5_reference_array(vals)
    scal = np.asarray([pack_rgb9e5_reference(v) for v in vals], dtype=np.uint32)
    np.testing.assert_array_equal(vec, scal)


def _channel_sweep(step: int) -> np.ndarray:
    bits = np.arange(0, 1 << 32, step, dtype=np.uint64).astype(np.uint32)
    return bits.view(np.float32)


def test_rgb9e5_single_channel_sweep():
    # every 997th float32 bit pattern in each channel (others 0), plus the
    # ties x = k + 0.5 - 2^-25 that round differently in float32
    x = _channel_sweep(997)
    for ch in range(3):
        v = np.zeros((len(x), 3), dtype=np.float32)
        v[:, ch] = x
        np.testing.assert_array_equal(oracle.pack_rgb9e5(v), pack_rgb9e5_reference_array(v))


def test_rgb9e5_small_channel_rounding_ties():
    # max channel 256 (scale 1): a small channel of 0.5 - 2^-25 must round down
    v = np.asarray([[256.0, 0.5 - 2.0**-25, 0.5], [300.0, 1.5 - 2.0**-23, 2.5]], dtype=np.float32)
    np.testing.assert_array_equal(oracle.pack_rgb9e5(v), pack_rgb9e5_reference_array(v))
    rng = np.random.default_rng(2)
    m = rng.uniform(1, 60000, size=20000).astype(np.float32)
    small = (rng.integers(0, 512, size=20000) + 0.5).astype(np.float64)
    scale = np.ldexp(1.0, np.floor(np.log2(m.astype(np.float64))).astype(int) + 1 - 9)
    ties = (small * scale).astype(np.float32)
    below = np.nextafter(ties, np.float32(0))
    v = np.stack([m, ties, below], axis=1).astype(np.float32)
    np.testing.assert_array_equal(oracle.pack_rgb9e5(v), pack_rgb9e5_reference_array(v))


def test_rgb9e5_quantization_bound():
    # test_mesh_colors_rgb9e5.py:130-146
    for value in [[0.1, 0.2, 0.3], [1.0, 2.0, 3.0], [10.0, 100.0, 1000.0], [0.001, 0.5, 12.0]]:
        packed = int(oracle.pack_rgb9e5(np.asarray([value], dtype=np.float32))[0])
        decoded = oracle.unpack_rgb9e5(packed)
        step = 2.0 ** ((packed >> 27) - 24)
        assert np.all(np.abs(decoded - np.asarray(value)) < step * 0.501)


# ---------------------------------------------------------------- raygen
def test_raygen_accepted_samples_map_back_to_their_texel(cornell):
    node, layout, arrays = cornell
    out, _ = oracle.texel_raygen(layout.face_table, layout.instance_table, 3, 1)
    b1 = out[:, 2].view(np.float32)
    b2 = out[:, 3].view(np.float32)
    b0 = 1.0 - b1 - b2
    tf = layout.texel_face_map()
    prefix = layout.front_prefix()
    fi = layout.face_instance_map()
    # the sample lies inside its triangle and within one lattice cell of its texel
    assert np.all(b0 >= -1e-6) and np.all(b1 >= 0) and np.all(b2 >= 0)
    res = layout.face_table[tf, 1].astype(np.float64)
    local = np.arange(layout.total_surface_texels) - prefix[fi[tf]].astype(np.int64) - layout.face_table[tf, 0]
    i = np.floor((2 * res + 3 - np.sqrt((2 * res + 3) ** 2 - 8 * local)) / 2)
    assert np.all(np.abs(b0 * res - i) <= 1.0 + 1e-4)
    # instance/prim agree with the face map
    np.testing.assert_array_equal(out[:, 0], fi[tf])
    np.testing.assert_array_equal(out[:, 1], tf - layout.instance_table[fi[tf], 0])


# ---------------------------------------------------------------- hits
def _random_rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, size=(n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8))
    rays[:, 0:3] = o
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = 1000.0
    return rays.astype(np.float32)


@pytest.mark.parametrize("which", ["cornell", "batch"])
def test_fp32_hits_match_fp64_brute_force(which):
    rng = np.random.default_rng(5)
    if which == "cornell":
        arrays = flatten_scene(cornell_box())
        rays = _random_rays(rng, 3000, [-0.95, 0.05, -0.95], [0.95, 1.95, 0.95])
    else:
        arrays = flatten_scene(batch_scene(11, 1500))
        rays = _random_rays(rng, 1500, [-8, 0, -8], [8, 4, 8])
    sc = oracle.OracleScene(arrays, use_bvh=True)
    t, b, inst, prim, occ = sc.intersect(rays)
    t64, b64, inst64, prim64, gap, edge = sc.intersect_fp64(rays.astype(np.float64))
    hit = inst != 0xFFFFFFFF
    hit64 = inst64 != 0xFFFFFFFF
    # rays whose nearest hit is clear of edges and of other hits are unambiguous
    clear = (~hit64) | ((gap > 1e-4 * np.maximum(t64, 1e-3)) & (edge > 1e-5))
    assert clear.mean() > 0.8
    np.testing.assert_array_equal(hit[clear], hit64[clear])
    both = clear & hit
    np.testing.assert_array_equal(inst[both], inst64[both])
    np.testing.assert_array_equal(prim[both], prim64[both])
    np.testing.assert_allclose(t[both], t64[both], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(b[both], b64[both], rtol=1e-5, atol=1e-5)
    # occlusion is exactly "there is a hit in [tmin, tmax]"
    np.testing.assert_array_equal(occ.astype(bool), hit)
    # the BVH and brute force give identical answers, ties included
    t2, b2, inst2, prim2, _ = oracle.OracleScene(arrays, use_bvh=False).intersect(rays)
    np.testing.assert_array_equal(t2.view(np.uint32), t.view(np.uint32))
    np.testing.assert_array_equal(inst2, inst)
    np.testing.assert_array_equal(prim2, prim)


# ---------------------------------------------------------------- tracer invariants
def _ground_scene(lighting: Lighting, albedo=1.0, double_sided=True) -> SceneNode:
    node = SceneNode()
    node.lighting = lighting
    mat = node.add_material(Material(base_color=(albedo,) * 3, double_sided=double_sided))
    mesh = node.add_mesh(Mesh.create_quad((4.0, 4.0)))
    node.add_instance(mesh, mat, node.add_transform(Transform()))
    return node


def _trace(node, spp=1, bounces=3, iteration=0, threads=4, tpu=2.0):
    layout = MeshColorsLayout.build(node, texels_per_unit=tpu, min_resolution=1, max_resolution=8)
    sc = oracle.OracleScene(flatten_scene(node))
    payload = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    cnt = sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, payload, iteration, spp,
                          bounces, True, threads=threads)
    return layout, payload, cnt


def test_sky_furnace_gives_pi_times_sky():
    # a lone plane under a constant sky S sees only sky: E = pi * S on both sides
    S = 0.7
    node = _ground_scene(Lighting(sun_intensity=0.0, sky_radiance=(S, S, S)))
    layout, payload, cnt = _trace(node)
    np.testing.assert_allclose(payload[:, :3], np.pi * S, rtol=2e-6)
    assert cnt[3] == layout.total_payload_count
    assert cnt[1] == 0  # no sun, no shadow rays


def test_sun_only_gives_lambert_cosine():
    # sun at 40 degrees elevation, no sky: E = I * n.l (cone jitter only)
    el = math.radians(40.0)
    node = _ground_scene(Lighting(sun_direction=(0.0, math.sin(el), math.cos(el)), sun_intensity=2.0,
                                  sky_radiance=(0.0, 0.0, 0.0)), double_sided=False)
    layout, payload, cnt = _trace(node)
    # each sample is within the sun cone (half-angle 0.268 deg); the cone
    # average of n.l is sin(el) to second order
    half = math.acos(Lighting().sun_cos_half_angle)
    np.testing.assert_allclose(payload[:, :3], 2.0 * math.sin(el), rtol=1.01 * half / math.tan(el))
    assert abs(payload[:, 0].mean() / (2.0 * math.sin(el)) - 1.0) < 2e-4
    assert cnt[1] == layout.total_surface_texels  # one unoccluded shadow ray per texel


def test_emitter_plane_direct_irradiance_converges():
    # receiver facing a large emitting plane at height h: E -> pi * Le * F,
    # F = view factor of a centered square of half-size a: closed form below
    Le, h, a = 2.0, 0.5, 20.0
    node = SceneNode()
    node.lighting = Lighting.black()
    recv = node.add_material(Material(base_color=(0.5,) * 3))
    emit = node.add_material(Material(base_color=(0.0,) * 3, emissive=(Le, Le, Le)))
    node.add_instance(node.add_mesh(Mesh.create_quad((0.25, 0.25))), recv, node.add_transform(Transform()))
    node.add_instance(node.add_mesh(Mesh.create_quad((2 * a, 2 * a))), emit,
                      node.add_transform(Transform((0, h, 0), (math.pi, 0, 0))))
    layout = MeshColorsLayout.build(node, texels_per_unit=16.0, min_resolution=1, max_resolution=4)
    sc = oracle.OracleScene(flatten_scene(node))
    payload = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    for it in range(64):
        sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, payload, it, 1, 1, it == 0,
                        threads=4)
    recv_slots = layout.instance_table[0, 3]
    # differential element under the centre of a parallel square of half-size a
    # at height h: four corner rectangles, F = (4/pi) s atan(s), s = x/sqrt(1+x^2)
    x = a / h
    s = x / math.sqrt(1 + x * x)
    F = (4 / math.pi) * s * math.atan(s)
    E = payload[:recv_slots, 0].astype(np.float64)
    assert abs(E.mean() - math.pi * Le * F) < 0.02 * math.pi * Le


def test_oracle_bvh_equals_brute_force_payload(cornell):
    node, layout, arrays = cornell
    outs = []
    for bvh in (False, True):
        sc = oracle.OracleScene(arrays, use_bvh=bvh)
        p = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        for it in range(2):
            sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, it, 1, 3, it == 0,
                            threads=4)
        outs.append(p)
    np.testing.assert_array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))


def test_running_mean_and_count_saturation(cornell):
    node, layout, arrays = cornell
    sc = oracle.OracleScene(arrays)
    p = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    work = np.arange(0, layout.total_surface_texels, 1000, dtype=np.uint64)
    sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, 0, 1, 3, True, work=work)
    first = p.copy()
    counts = p[:, 3].view(np.uint32)
    touched = counts == 1
    assert touched.sum() >= len(work)
    # count saturates at 65535 and the mean weight then stays 1/65535
    p[:, 3] = np.uint32(65535).view(np.float32)
    sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, 1, 1, 3, False, work=work)
    assert np.all(p[touched, 3].view(np.uint32) == 65535)
    # reset restarts the mean from this iteration's sample
    sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, 0, 1, 3, True, work=work)
    np.testing.assert_array_equal(p[touched].view(np.uint32), first[touched].view(np.uint32))


def test_lattice_point_resolve_reads_its_texel(cornell):
    # a constant cache resolves to emissive + albedo/pi * constant
    node, layout, arrays = cornell
    sc = oracle.OracleScene(arrays)
    payload = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
    payload[:, :3] = 2.0
    cam = node.camera
    cam.width, cam.height = 64, 64
    cam.recompute()
    img = sc.resolve(layout.face_table, layout.instance_table, layout.side_table, cam.pack(), 64, 64,
                     payload=payload)
    center = img[40, 32, :3]
    np.testing.assert_allclose(center, np.asarray([0.73] * 3) / np.pi * 2.0, rtol=1e-5)


def test_oracle_bvh_queries_equal_brute_force_sunlit_scene():
    # every extend and shadow query of a sunlit batch scene, repeated by brute
    # force inside the oracle: the box-test slack keeps culling exact even for
    # the coplanar box bottoms buried in the terrain
    node = batch_scene(21, 6000)
    layout = MeshColorsLayout.build(node, texels_per_unit=8.0, min_resolution=1, max_resolution=8,
                                    max_total_texels=2_000_000)
    lib = oracle.liboracle()
    lib.orc_set_crosscheck(1)
    try:
        sc = oracle.OracleScene(flatten_scene(node))
        p = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, p, 0, 1, 3, True,
                        threads=oracle.cpu_count())
        ray = np.zeros(8, dtype=np.float32)
        kind = lib.orc_crosscheck_result(ray.ctypes.data_as(oracle.f32p))
    finally:
        lib.orc_set_crosscheck(0)
    assert kind == 0, f"BVH and brute force disagree on ray {ray.tolist()} (kind {kind})"
