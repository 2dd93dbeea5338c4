"""Pins of the oracle's texture sampling, normal mapping and alpha test
(scene.slang:124-149, 313-368; common.slang:516-577; texture_manager.py:
93-94) by what fixes them independently of its code:

* the texel decode table against the IEC 61966-2-1 sRGB definition;
* bilinear wrap sampling against texel centres, an independent float64
  restatement of the D3D filtering rule, integer-period wrap, the exact
  reproduction of a linear texel field, and decode-before-filter;
* the murmur mix against scikit-learn's MurmurHash3_x86_32 (the reference
  omits the length word, undone here through the finaliser's inverse), and
  the voxel key's periods and truncation toward zero;
* mask and blend alpha tests on rays through a textured card in front of a
  backstop: hits where the texture is opaque, >= at the cutoff, the blend
  hit rate equal to the opacity, occlusion equal to the closest hit;
* normal mapping: sun-only irradiance on a quad equals I * n.l with n the
  tangent-space normal taken to the world through the quad's uv frame.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from nan_amd.material import ALPHA_MODE_BLEND, ALPHA_MODE_MASK, Material
from nan_amd.mesh import Mesh
from nan_amd.mesh_colors import MeshColorsLayout
from nan_amd.scene_arrays import flatten_scene
from nan_amd.scene_node import Lighting, SceneNode
from nan_amd.transform import Transform


def _srgb_decode(c: np.ndarray) -> np.ndarray:
    x = c / 255.0
    return np.where(x <= 0.04045, x / 12.92, ((x + 0.055) / 1.055) ** 2.4)


def test_texel_lut_is_unorm_then_srgb():
    lut = oracle.texel_lut()
    c = np.arange(256, dtype=np.float64)
    np.testing.assert_array_equal(lut[:256], (c / 255.0).astype(np.float32))
    np.testing.assert_array_equal(lut[256:], _srgb_decode(c).astype(np.float32))
    assert lut[256] == 0.0 and lut[511] == 1.0 and lut[255] == 1.0
    assert np.all(np.diff(lut[256:]) > 0)
    assert abs(lut[256 + 128] - 0.2158605) < 1e-6  # sRGB 128 is 21.6% linear


def _texture_scene(images, srgb_slots=("base_color_texture",)) -> oracle.OracleScene:
    """A one-quad scene whose materials carry ``images`` (one material per
    image, in the slot named for it: base colour is sRGB, normal is not)."""
    node = SceneNode()
    mesh = node.add_mesh(Mesh.create_quad((1.0, 1.0)))
    xf = node.add_transform(Transform())
    for k, img in enumerate(images):
        slot = srgb_slots[k % len(srgb_slots)]
        mat = node.add_material(Material(**{slot: img}))
        node.add_instance(mesh, mat, xf)
    return oracle.OracleScene(flatten_scene(node))


def _bilinear_f64(img_lin: np.ndarray, uv: np.ndarray) -> np.ndarray:
    """D3D/Vulkan linear filtering with wrap, written independently: texel
    centres at (i + 0.5) / n, weights in float64, indices modulo n."""
    h, w = img_lin.shape[:2]
    x = uv[:, 0] * w - 0.5
    y = uv[:, 1] * h - 0.5
    x0 = np.floor(x)
    y0 = np.floor(y)
    fx = (x - x0)[:, None]
    fy = (y - y0)[:, None]
    i0 = x0.astype(np.int64) % w
    j0 = y0.astype(np.int64) % h
    i1 = (i0 + 1) % w
    j1 = (j0 + 1) % h
    return ((1 - fx) * (1 - fy) * img_lin[j0, i0] + fx * (1 - fy) * img_lin[j0, i1] +
            (1 - fx) * fy * img_lin[j1, i0] + fx * fy * img_lin[j1, i1])


@pytest.fixture(scope="module")
def textures():
    rng = np.random.default_rng(5)
    imgs = [rng.integers(0, 256, size=(5, 7, 4), dtype=np.uint8),    # base colour: sRGB rgb
            rng.integers(0, 256, size=(3, 11, 4), dtype=np.uint8)]   # normal slot: linear
    sc = _texture_scene(imgs, ("base_color_texture", "normal_texture"))
    lut = oracle.texel_lut().astype(np.float64)
    lin = []
    for img, srgb in zip(imgs, (True, False)):
        f = lut[img.astype(np.int64)]
        if srgb:
            f[:, :, :3] = lut[256 + img[:, :, :3].astype(np.int64)]
        lin.append(f)
    return sc, imgs, lin


def test_texel_centres_read_their_texel(textures):
    sc, imgs, lin = textures
    for k, img in enumerate(imgs):
        h, w = img.shape[:2]
        jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
        uv = np.stack([(ii.ravel() + 0.5) / w, (jj.ravel() + 0.5) / h], axis=1)
        got = sc.sample_texture(k, uv)
        np.testing.assert_allclose(got, lin[k][jj.ravel(), ii.ravel()], atol=2e-6)


def test_bilinear_wrap_matches_float64_filtering(textures):
    sc, imgs, lin = textures
    rng = np.random.default_rng(6)
    uv = rng.uniform(-3.0, 3.0, size=(4000, 2))
    for k in range(len(imgs)):
        got = sc.sample_texture(k, uv.astype(np.float32))
        ref = _bilinear_f64(lin[k], uv.astype(np.float32).astype(np.float64))
        np.testing.assert_allclose(got, ref, atol=2e-5)


def test_integer_shifts_of_uv_sample_the_same(textures):
    sc, imgs, _ = textures
    rng = np.random.default_rng(7)
    uv = rng.integers(0, 256, size=(500, 2)) / 256.0  # exact in float32 with the shifts below
    base = sc.sample_texture(0, uv)
    for shift in ((1, 0), (0, 1), (-3, 2), (5, -4)):
        np.testing.assert_array_equal(sc.sample_texture(0, uv + np.array(shift)), base)


def test_linear_texel_field_is_reproduced():
    # unorm texels 9*i + 20*j: the bilinear sample inside the texture is the
    # linear function at (x, y) = (u*w - 0.5, v*h - 0.5), exactly up to rounding
    w, h = 8, 6
    jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    img = np.zeros((h, w, 4), dtype=np.uint8)
    img[:, :, 0] = 9 * ii + 20 * jj
    img[:, :, 1] = 255 - img[:, :, 0]
    img[:, :, 3] = 255
    sc = _texture_scene([img], ("normal_texture",))
    rng = np.random.default_rng(8)
    x = rng.uniform(0.0, w - 1.0, size=3000)
    y = rng.uniform(0.0, h - 1.0, size=3000)
    uv = np.stack([(x + 0.5) / w, (y + 0.5) / h], axis=1).astype(np.float32)
    got = sc.sample_texture(0, uv)
    xs = uv[:, 0].astype(np.float64) * w - 0.5
    ys = uv[:, 1].astype(np.float64) * h - 0.5
    np.testing.assert_allclose(got[:, 0], (9 * xs + 20 * ys) / 255.0, atol=3e-6)
    np.testing.assert_allclose(got[:, 1], 1.0 - (9 * xs + 20 * ys) / 255.0, atol=3e-6)


def test_srgb_is_decoded_before_filtering():
    img = np.array([[[0, 0, 0, 0], [255, 255, 255, 255]]], dtype=np.uint8)  # 2x1
    sc = _texture_scene([img])
    got = sc.sample_texture(0, np.array([[0.5, 0.5]], dtype=np.float32))[0]
    # halfway between the texel centres: mean of the decoded values (0.5), not
    # the decode of the mean byte (0.21); alpha is never sRGB
    np.testing.assert_allclose(got, [0.5, 0.5, 0.5, 0.5], atol=1e-7)


# ------------------------------------------------------------------ hashing
def _fmix32_inverse(h: int) -> int:
    m = 0xFFFFFFFF
    h ^= h >> 16
    h = (h * pow(0xC2B2AE35, -1, 1 << 32)) & m
    h ^= (h >> 13) ^ (h >> 26)
    h = (h * pow(0x85EBCA6B, -1, 1 << 32)) & m
    h ^= h >> 16
    return h


def test_murmur_key_mix_is_murmurhash3_without_the_length_word():
    from sklearn.utils import murmurhash3_32

    lib = oracle.liboracle()
    rng = np.random.default_rng(9)
    for h, k in rng.integers(0, 1 << 32, size=(300, 2), dtype=np.uint64).tolist() + [[0, 0], [0xFFFFFFFF, 1]]:
        ours = int(lib.orc_murmur_key_mix(h, k))
        key = int(np.array([k], dtype=np.uint32).view(np.int32)[0])
        theirs = int(murmurhash3_32(key, seed=int(h), positive=True))
        # MurmurHash3_x86_32 xors the key length (4) into the state before its finaliser
        assert _fmix32_inverse(ours) == _fmix32_inverse(theirs) ^ 4


def test_position_hash_voxels_periods_and_truncation():
    def h(p, frame=3):
        return oracle.alpha_hash(np.asarray(p, dtype=np.float32).reshape(-1, 3), frame)

    base = np.array([[0.12343, 0.45673, 0.78903]], dtype=np.float32)
    v = 1e-4
    # periods of the key: x and z keep 15 bits of the voxel, y 12 bits
    for shift in ((32768 * v, 0, 0), (0, 4096 * v, 0), (0, 0, 32768 * v)):
        assert h(base + np.float32(shift)) == h(base)
    # one voxel, or half a period, along any axis changes the hash
    for shift in ((v, 0, 0), (0, v, 0), (0, 0, v), (16384 * v, 0, 0), (0, 2048 * v, 0), (0, 0, 16384 * v)):
        assert h(base + np.float32(shift)) != h(base)
    # the same voxel hashes the same; the frame index changes it
    assert h(base + np.float32(0.3 * v)) == h(base)
    assert h(base, 4) != h(base, 3)
    # int3() truncates toward zero: (-0.5, 0.5) voxels share voxel 0
    assert h([[-0.5 * v, 0.2, 0.2]]) == h([[0.5 * v, 0.2, 0.2]])
    assert h([[-1.5 * v, 0.2, 0.2]]) != h([[1.5 * v, 0.2, 0.2]])


# ------------------------------------------------------------------ alpha test
def _card_scene(material: Material) -> oracle.OracleScene:
    """A 2x2 card in the z = 0 plane (u = (x + 1) / 2, v = (y + 1) / 2)
    with ``material``, and an opaque backstop at z = -1."""
    node = SceneNode()
    card = node.add_mesh(Mesh.create_quad_xy((2.0, 2.0)))
    back = node.add_mesh(Mesh.create_quad_xy((8.0, 8.0)))
    node.add_instance(card, node.add_material(material), node.add_transform(Transform()))
    node.add_instance(back, node.add_material(Material()), node.add_transform(Transform((0.0, 0.0, -1.0))))
    return oracle.OracleScene(flatten_scene(node))


def _down_rays(xy: np.ndarray, tmax: float = 1000.0) -> np.ndarray:
    r = np.zeros((len(xy), 8), dtype=np.float32)
    r[:, 0:2] = xy
    r[:, 2] = 5.0
    r[:, 6] = -1.0
    r[:, 7] = tmax
    return r


def _half_opaque_texture() -> np.ndarray:
    img = np.full((4, 4, 4), 200, dtype=np.uint8)
    img[:, :2, 3] = 0    # u < 0.5: transparent
    img[:, 2:, 3] = 255  # u > 0.5: opaque
    return img


def test_mask_keeps_the_opaque_half_and_ties_at_the_cutoff():
    sc = _card_scene(Material(base_color_texture=_half_opaque_texture(), alpha_mode=ALPHA_MODE_MASK,
                              alpha_cutoff=0.5))
    us = np.array([0.125, 0.375, 0.625, 0.875, 0.5])
    xy = np.stack([2 * us - 1, np.zeros_like(us)], axis=1)
    t, _, inst, _, occ = sc.intersect(_down_rays(xy))
    # texel centres read 0 or 1; at u = 0.5 the filtered alpha is exactly 0.5 >= cutoff
    np.testing.assert_array_equal(inst, [1, 1, 0, 0, 0])
    np.testing.assert_array_equal(t, np.float32([6, 6, 5, 5, 5]))
    assert occ.all()  # the backstop occludes everything
    _, _, inst2, _, occ2 = sc.intersect(_down_rays(xy, tmax=5.5))
    np.testing.assert_array_equal(occ2, (inst2 == 0).astype(np.uint32))
    np.testing.assert_array_equal(occ2, [0, 0, 1, 1, 1])


def test_opaque_material_ignores_texture_alpha():
    sc = _card_scene(Material(base_color_texture=_half_opaque_texture()))
    xy = np.array([[-0.75, 0.0], [0.75, 0.0]])
    _, _, inst, _, _ = sc.intersect(_down_rays(xy))
    np.testing.assert_array_equal(inst, [0, 0])


@pytest.mark.parametrize("cutoff,kept", [(0.5, True), (0.502, False), (1.5, False)])
def test_mask_cutoff_against_uniform_alpha(cutoff, kept):
    img = np.full((2, 2, 4), 128, dtype=np.uint8)  # alpha 128/255 = 0.50196
    sc = _card_scene(Material(base_color_texture=img, alpha_mode=ALPHA_MODE_MASK, alpha_cutoff=cutoff))
    _, _, inst, _, _ = sc.intersect(_down_rays(np.array([[0.1, 0.2]])))
    assert (inst[0] == 0) == kept


def test_mask_without_texture_has_opacity_one():
    for cutoff, kept in ((1.0, True), (1.01, False)):
        sc = _card_scene(Material(alpha_mode=ALPHA_MODE_MASK, alpha_cutoff=cutoff))
        _, _, inst, _, _ = sc.intersect(_down_rays(np.array([[0.1, 0.2]])))
        assert (inst[0] == 0) == kept


def test_blend_keeps_hits_at_the_opacity_rate():
    img = np.full((3, 3, 4), 120, dtype=np.uint8)
    img[:, :, 3] = 77  # opacity 0.30196
    sc = _card_scene(Material(base_color_texture=img, alpha_mode=ALPHA_MODE_BLEND))
    rng = np.random.default_rng(10)
    xy = rng.uniform(-0.99, 0.99, size=(20000, 2))
    rates, kept = [], []
    for frame in (0, 1, 2):
        sc.set_frame_index(frame)
        _, _, inst, _, _ = sc.intersect(_down_rays(xy))
        k = inst == 0
        kept.append(k)
        rates.append(k.mean())
        _, _, inst2, _, occ = sc.intersect(_down_rays(xy, tmax=5.5))
        np.testing.assert_array_equal(occ.astype(bool), k)  # any-hit agrees with closest hit
        np.testing.assert_array_equal(inst2 == 0, k)
    p = 77 / 255
    sigma = math.sqrt(p * (1 - p) / len(xy))
    for r in rates:
        assert abs(r - p) < 4.5 * sigma, (r, p)
    # a new frame index redraws each hit independently: about 2p(1-p) of them flip
    flipped = (kept[0] != kept[1]).mean()
    assert abs(flipped - 2 * p * (1 - p)) < 0.02
    # the same frame index gives the same hits
    sc.set_frame_index(1)
    _, _, inst, _, _ = sc.intersect(_down_rays(xy))
    np.testing.assert_array_equal(inst == 0, kept[1])


def test_blend_without_texture_is_opaque():
    sc = _card_scene(Material(alpha_mode=ALPHA_MODE_BLEND))
    xy = np.random.default_rng(12).uniform(-0.9, 0.9, size=(500, 2))
    _, _, inst, _, _ = sc.intersect(_down_rays(xy))
    assert (inst == 0).all()


# ------------------------------------------------------------------ normal mapping
def _normal_mapped_quad(texel, sun_dir, rotation=(0.0, 0.0, 0.0), scale=1.0) -> tuple:
    node = SceneNode()
    node.lighting = Lighting(sun_direction=tuple(sun_dir), sun_intensity=2.0, sky_radiance=(0.0, 0.0, 0.0),
                             sun_cos_half_angle=math.cos(1e-5))
    img = np.zeros((2, 2, 4), dtype=np.uint8)
    img[:, :, :3] = texel
    img[:, :, 3] = 255
    # black: a bounce ray sampled about the tilted normal may re-hit the quad,
    # and then brings back nothing
    mat = node.add_material(Material(base_color=(0.0, 0.0, 0.0), normal_texture=img))
    node.add_instance(node.add_mesh(Mesh.create_quad((2.0, 2.0))), mat,
                      node.add_transform(Transform((0.0, 0.0, 0.0), rotation, (scale, scale, scale))))
    return node, Transform((0.0, 0.0, 0.0), rotation, (scale, scale, scale)).matrix


@pytest.mark.parametrize("texel,rotation", [((200, 128, 220), (0.0, 0.0, 0.0)), ((60, 190, 230), (0.0, 0.0, 0.0)),
                                            ((200, 90, 235), (30.0, 50.0, 10.0))])
def test_normal_map_tilts_the_lambert_cosine(texel, rotation):
    # create_quad: u runs along +x, v along +z, normal +y.  The tangent frame
    # of scene.slang:346-366 is then T = +x and B = cross(N, T) = -z, so the
    # tangent-space normal (a, b, c) is a*x - b*z + c*y, rotated with the
    # instance (uniform scale: the inverse transpose is the rotation)
    t = np.asarray(texel, dtype=np.float64) / 255.0 * 2.0 - 1.0
    t /= np.linalg.norm(t)
    n_obj = np.array([t[0], t[2], -t[1]])
    for sun in ((0.3, 0.8, 0.2), (-0.5, 0.6, 0.3), (0.1, 0.9, -0.6)):
        sun = np.asarray(sun) / np.linalg.norm(sun)
        node, m = _normal_mapped_quad(texel, sun, rotation)
        rot = np.asarray(m, dtype=np.float64)[:3, :3]
        n_world = rot @ n_obj
        n_world /= np.linalg.norm(n_world)
        flat = rot @ np.array([0.0, 1.0, 0.0])
        flat /= np.linalg.norm(flat)
        sun_w = sun  # the sun direction is given in world space
        if flat @ sun_w <= 0.05 or n_world @ sun_w <= 0.05:
            continue
        layout = MeshColorsLayout.build(node, texels_per_unit=2.0, min_resolution=1, max_resolution=4)
        sc = oracle.OracleScene(flatten_scene(node))
        payload = np.zeros((layout.total_payload_count, 4), dtype=np.float32)
        sc.trace_texels(layout.face_table, layout.instance_table, layout.side_table, payload, 0, 1, 1, True,
                        threads=4)
        e = payload[: layout.total_surface_texels, :3]
        np.testing.assert_allclose(e, 2.0 * (n_world @ sun_w), rtol=2e-4)
        assert abs(2.0 * (flat @ sun_w) - e[0, 0]) > 0.01  # the map changed the answer
