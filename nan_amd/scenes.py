"""Seeded procedural scenes for the texture-space workloads.

* :func:`cornell_box` - the small Cornell box: 2-unit cube, 5 double-sided
  walls (10 triangles), axis-aligned tall (0.5x1.2x0.5) and short
  (0.5x0.6x0.5) blocks (24 triangles), a 0.5x0.4 ceiling light with
  emission (17, 12, 4); no sun and no sky.  The instance order follows the
  reference's demo (scene_node.py:721-830).
* :func:`synthetic_scene` - the large scene: a 512x512-quad value-noise
  heightfield (524,288 triangles), a 64x64 grid of extruded boxes with
  lognormal heights (49,152 triangles) and 500 level-3 icospheres
  (640,000 triangles) with radii in U[0.2, 1.5]; per-object albedo in
  U[0.1, 0.8], 5% double-sided materials, a sun at 35 degrees elevation
  and a constant sky of 0.3.  ``scale`` multiplies every object count
  (the 4M-triangle scene uses ``scale=3.3``).
* :func:`batch_scene` - a small seeded mix of the same primitives with a
  requested triangle count, for many-scene batches.
* :func:`textured_scene` - the batch mix with textured materials: a noise
  base-colour texture and a bump normal map on the terrain, checker boxes,
  alpha-masked double-sided cards with uvs that wrap, blend panes, spheres
  with emissive textures, and texture sizes from 1x1 to non-powers of two.
"""

from __future__ import annotations

import math

import numpy as np

from .material import Material
from .mesh import Mesh
from .scene_node import Lighting, SceneNode
from .transform import Transform


# ---------------------------------------------------------------------------
# Cornell box
# ---------------------------------------------------------------------------
def cornell_box() -> SceneNode:
    node = SceneNode()
    node.camera.target = np.array([0.0, 1.0, 0.0])
    node.camera.position = np.array([0.0, 1.0, 2.95])
    node.camera.fov = 60.0
    node.camera.recompute()
    node.lighting = Lighting.black()
    white = node.add_material(Material(base_color=(0.73, 0.73, 0.73), double_sided=True))
    red = node.add_material(Material(base_color=(0.65, 0.05, 0.05), double_sided=True))
    green = node.add_material(Material(base_color=(0.12, 0.45, 0.15), double_sided=True))
    white_block = node.add_material(Material(base_color=(0.73, 0.73, 0.73)))
    light = node.add_material(Material(base_color=(1.0, 1.0, 1.0), emissive=(17.0, 12.0, 4.0)))
    quad = node.add_mesh(Mesh.create_quad((2, 2)))
    back = node.add_mesh(Mesh.create_quad_xy((2, 2), face_positive_z=True))
    left = node.add_mesh(Mesh.create_quad_yz((2, 2), face_positive_x=True))
    right = node.add_mesh(Mesh.create_quad_yz((2, 2), face_positive_x=False))
    node.add_instance(quad, white, node.add_transform(Transform((0, 0, 0))))
    node.add_instance(quad, white, node.add_transform(Transform((0, 2, 0), (math.pi, 0, 0))))
    node.add_instance(back, white, node.add_transform(Transform((0, 1, -1))))
    node.add_instance(left, red, node.add_transform(Transform((-1, 1, 0))))
    node.add_instance(right, green, node.add_transform(Transform((1, 1, 0))))
    light_mesh = node.add_mesh(Mesh.create_quad((0.5, 0.4)))
    node.add_instance(light_mesh, light, node.add_transform(Transform((0, 1.99, 0), (math.pi, 0, 0))))
    tall = node.add_mesh(Mesh.create_cube((0.5, 1.2, 0.5)))
    node.add_instance(tall, white_block, node.add_transform(Transform((-0.35, 0.6, -0.3))))
    short = node.add_mesh(Mesh.create_cube((0.5, 0.6, 0.5)))
    node.add_instance(short, white_block, node.add_transform(Transform((0.35, 0.3, 0.3))))
    return node


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
def _value_noise(rng: np.random.Generator, xs: np.ndarray, ys: np.ndarray, octaves: int) -> np.ndarray:
    """Sum of ``octaves`` bilinear value-noise layers with smoothstep weights,
    frequency doubling and amplitude halving; result in about [-1, 1]."""
    out = np.zeros_like(xs, dtype=np.float64)
    amp = 1.0
    total = 0.0
    freq = 1.0
    for _ in range(octaves):
        size = int(math.ceil(freq)) + 2
        lattice = rng.uniform(-1.0, 1.0, size=(size + 1, size + 1))
        u = xs * freq
        v = ys * freq
        i = np.floor(u).astype(np.int64)
        j = np.floor(v).astype(np.int64)
        fu = u - i
        fv = v - j
        su = fu * fu * (3 - 2 * fu)
        sv = fv * fv * (3 - 2 * fv)
        i = np.clip(i, 0, size - 1)
        j = np.clip(j, 0, size - 1)
        a = lattice[i, j]
        b = lattice[i + 1, j]
        c = lattice[i, j + 1]
        d = lattice[i + 1, j + 1]
        out += amp * ((a * (1 - su) + b * su) * (1 - sv) + (c * (1 - su) + d * su) * sv)
        total += amp
        amp *= 0.5
        freq *= 2.0
    return out / total


def heightfield_mesh(rng: np.random.Generator, quads: int, extent: float, amplitude: float,
                     base_freq: float = 4.0) -> Mesh:
    """``quads x quads`` grid over ``[-extent/2, extent/2]^2`` in XZ with
    4-octave value-noise heights; smooth per-vertex normals."""
    n = quads + 1
    lin = np.linspace(-0.5 * extent, 0.5 * extent, n)
    gx, gz = np.meshgrid(lin, lin, indexing="ij")
    nx = (gx / extent + 0.5) * base_freq
    nz = (gz / extent + 0.5) * base_freq
    h = amplitude * _value_noise(rng, nx, nz, 4)
    pos = np.stack([gx, h, gz], axis=-1)
    # central-difference normals
    dhx = np.gradient(h, lin, axis=0)
    dhz = np.gradient(h, lin, axis=1)
    nrm = np.stack([-dhx, np.ones_like(h), -dhz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    uv = np.stack([gx / extent + 0.5, gz / extent + 0.5], axis=-1)
    verts = np.concatenate([pos, nrm, uv], axis=-1).reshape(-1, 8).astype(np.float32)
    idx = np.arange(n * n, dtype=np.int64).reshape(n, n)
    a = idx[:-1, :-1].reshape(-1)
    b = idx[1:, :-1].reshape(-1)
    c = idx[:-1, 1:].reshape(-1)
    d = idx[1:, 1:].reshape(-1)
    # counter-clockwise seen from +Y: (a, c, b) and (b, c, d)
    tris = np.stack([np.stack([a, c, b], 1), np.stack([b, c, d], 1)], axis=1).reshape(-1, 3)
    return Mesh(verts, tris.astype(np.uint32))


def icosphere_mesh(level: int) -> Mesh:
    """Unit icosphere subdivided ``level`` times (20 * 4^level triangles)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache: dict[tuple[int, int], int] = {}

        def midpoint(a: int, b: int) -> int:
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        new_faces = []
        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    p = np.stack(verts)
    v = np.zeros((len(p), 8), dtype=np.float32)
    v[:, 0:3] = p
    v[:, 3:6] = p
    v[:, 6] = 0.5 + np.arctan2(p[:, 2], p[:, 0]) / (2 * math.pi)
    v[:, 7] = 0.5 - np.arcsin(np.clip(p[:, 1], -1, 1)) / math.pi
    return Mesh(v, np.asarray(faces, dtype=np.uint32))


def box_mesh() -> Mesh:
    """Unit cube on the ground plane: x,z in [-0.5,0.5], y in [0,1]."""
    m = Mesh.create_cube((1.0, 1.0, 1.0))
    v = m.vertices.copy()
    v[:, 1] += np.float32(0.5)
    return Mesh(v, m.indices.copy())


def _terrain_height(mesh: Mesh, quads: int, extent: float, x: float, z: float) -> float:
    n = quads + 1
    i = int(round((x / extent + 0.5) * quads))
    j = int(round((z / extent + 0.5) * quads))
    i = min(max(i, 0), quads)
    j = min(max(j, 0), quads)
    return float(mesh.vertices[i * n + j, 1])


# ---------------------------------------------------------------------------
# large synthetic scene
# ---------------------------------------------------------------------------
def synthetic_scene(seed: int = 0, scale: float = 1.0, *, terrain_quads: int = 512, box_grid: int = 64,
                    spheres: int = 500, sphere_level: int = 3, extent: float = 64.0,
                    double_sided: float = 0.05, single_sided_terrain: bool = False) -> SceneNode:
    """Heightfield + extruded boxes + icospheres (see module docstring);
    ``scale`` multiplies the triangle count (terrain and box grid by its
    square root per side, spheres linearly) and ``double_sided`` is the
    fraction of object materials that are double-sided (the terrain's draw
    is overridden to single-sided with ``single_sided_terrain``)."""
    rng = np.random.default_rng(seed)
    node = SceneNode()
    node.lighting = Lighting.sun_and_sky(35.0, 30.0, sky=0.3)
    quads = int(round(terrain_quads * math.sqrt(scale)))
    grid = int(round(box_grid * math.sqrt(scale)))
    n_spheres = int(round(spheres * scale))
    terrain = heightfield_mesh(rng, quads, extent, 2.0)
    n_objects = 1 + grid * grid + n_spheres
    double_sided = rng.random(n_objects) < double_sided
    if single_sided_terrain:
        double_sided[0] = False
    albedo = rng.uniform(0.1, 0.8, size=(n_objects, 3))
    identity = node.add_transform(Transform())

    def material(k: int) -> int:
        return node.add_material(Material(base_color=tuple(albedo[k]), double_sided=bool(double_sided[k])))

    node.add_instance(node.add_mesh(terrain), material(0), identity)

    box = node.add_mesh(box_mesh())
    spacing = extent * 0.8 / grid
    heights = rng.lognormal(1.0, 0.6, size=grid * grid)
    heights = np.minimum(heights, 12.0)
    k = 1
    for gi in range(grid):
        for gj in range(grid):
            x = (gi - grid / 2 + 0.5) * spacing
            z = (gj - grid / 2 + 0.5) * spacing
            w = spacing * rng.uniform(0.3, 0.6)
            d = spacing * rng.uniform(0.3, 0.6)
            y = _terrain_height(terrain, quads, extent, x, z) - 0.2
            t = node.add_transform(Transform((x, y, z), (0.0, 0.0, 0.0), (w, heights[gi * grid + gj], d)))
            node.add_instance(box, material(k), t)
            k += 1

    sphere = node.add_mesh(icosphere_mesh(sphere_level))
    for _ in range(n_spheres):
        r = rng.uniform(0.2, 1.5)
        x, z = rng.uniform(-0.45 * extent, 0.45 * extent, size=2)
        y = _terrain_height(terrain, quads, extent, x, z) + r * rng.uniform(0.2, 1.0) + rng.uniform(0.0, 3.0)
        t = node.add_transform(Transform((x, y, z), (0.0, 0.0, 0.0), (r, r, r)))
        node.add_instance(sphere, material(k), t)
        k += 1

    node.camera.position = np.array([0.0, 18.0, 0.55 * extent])
    node.camera.target = np.array([0.0, 0.0, 0.0])
    node.camera.fov = 60.0
    node.camera.recompute()
    return node


def batch_scene(seed: int, triangles: int) -> SceneNode:
    """A seeded mix of heightfield, boxes and icospheres with about
    ``triangles`` triangles (at least one of each primitive)."""
    rng = np.random.default_rng(seed)
    node = SceneNode()
    node.lighting = Lighting.sun_and_sky(float(rng.uniform(20, 60)), float(rng.uniform(0, 360)), sky=0.3)
    extent = 16.0
    share = rng.dirichlet([2.0, 1.0, 1.0])
    quads = max(2, int(math.sqrt(max(2.0, share[0] * triangles / 2))))
    n_boxes = max(1, int(share[1] * triangles / 12))
    n_spheres = max(1, int(share[2] * triangles / 320))
    terrain = heightfield_mesh(rng, quads, extent, 1.0)
    identity = node.add_transform(Transform())
    n_objects = 1 + n_boxes + n_spheres
    double_sided = rng.random(n_objects) < 0.05
    albedo = rng.uniform(0.1, 0.8, size=(n_objects, 3))

    def material(k: int) -> int:
        return node.add_material(Material(base_color=tuple(albedo[k]), double_sided=bool(double_sided[k])))

    node.add_instance(node.add_mesh(terrain), material(0), identity)
    box = node.add_mesh(box_mesh())
    k = 1
    for _ in range(n_boxes):
        x, z = rng.uniform(-0.45 * extent, 0.45 * extent, size=2)
        s = rng.uniform(0.2, 0.8, size=2)
        h = float(min(rng.lognormal(0.0, 0.6), 4.0))
        y = _terrain_height(terrain, quads, extent, x, z) - 0.1
        node.add_instance(box, material(k), node.add_transform(Transform((x, y, z), (0, 0, 0), (s[0], h, s[1]))))
        k += 1
    sphere = node.add_mesh(icosphere_mesh(2))
    for _ in range(n_spheres):
        r = rng.uniform(0.2, 1.0)
        x, z = rng.uniform(-0.45 * extent, 0.45 * extent, size=2)
        y = _terrain_height(terrain, quads, extent, x, z) + r + rng.uniform(0.0, 1.5)
        node.add_instance(sphere, material(k), node.add_transform(Transform((x, y, z), (0, 0, 0), (r, r, r))))
        k += 1
    node.camera.position = np.array([0.0, 6.0, 0.6 * extent])
    node.camera.target = np.array([0.0, 0.0, 0.0])
    node.camera.recompute()
    return node


# ---------------------------------------------------------------------------
# textured scene
# ---------------------------------------------------------------------------
def _noise_texture(rng: np.random.Generator, w: int, h: int, alpha=None) -> np.ndarray:
    """RGBA8 value noise, alpha 255 unless given (a scalar or an array)."""
    ys, xs = np.meshgrid(np.linspace(0, 4, h), np.linspace(0, 4, w), indexing="ij")
    img = np.zeros((h, w, 4), dtype=np.uint8)
    for c in range(3):
        v = _value_noise(rng, xs, ys, 3)
        img[:, :, c] = np.clip(40 + 200 * v, 0, 255).astype(np.uint8)
    img[:, :, 3] = 255 if alpha is None else alpha
    return img


def _checker(w: int, h: int, cells: int, a, b) -> np.ndarray:
    ys, xs = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    on = ((xs * cells // w) + (ys * cells // h)) % 2 == 0
    img = np.where(on[:, :, None], np.asarray(a, dtype=np.uint8), np.asarray(b, dtype=np.uint8))
    return np.ascontiguousarray(img, dtype=np.uint8)


def _bump_normal_map(rng: np.random.Generator, w: int, h: int, strength: float) -> np.ndarray:
    """Tangent-space normal map (RGB = n * 0.5 + 0.5) of a noise height."""
    ys, xs = np.meshgrid(np.linspace(0, 3, h), np.linspace(0, 3, w), indexing="ij")
    height = _value_noise(rng, xs, ys, 2) * strength
    gy, gx = np.gradient(height)
    n = np.stack([-gx * w / 3, -gy * h / 3, np.ones_like(height)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    img = np.zeros((h, w, 4), dtype=np.uint8)
    img[:, :, :3] = np.clip(np.rint((n * 0.5 + 0.5) * 255), 0, 255).astype(np.uint8)
    img[:, :, 3] = 255
    return img


def _card_mesh(u_repeat: float) -> Mesh:
    """XY quad (two triangles) whose uvs run over [-u_repeat/2, u_repeat/2]
    x [0, 1], so a texture wraps across it."""
    m = Mesh.create_quad_xy((1.0, 1.0))
    v = m.vertices.copy()
    v[:, 6] = (v[:, 6] - 0.5) * np.float32(u_repeat)
    return Mesh(v, m.indices.copy())


def textured_scene(seed: int, triangles: int = 6000) -> SceneNode:
    """A seeded textured scene of about ``triangles`` triangles: the terrain
    of :func:`batch_scene` with a base-colour texture and a normal map,
    checker boxes (some with normal maps), alpha-masked double-sided cards,
    blend panes (textured and untextured), spheres with emissive textures,
    an untextured blend material (always kept) and a mask material whose
    cutoff exceeds 1 (never kept)."""
    from .material import ALPHA_MODE_BLEND, ALPHA_MODE_MASK

    rng = np.random.default_rng(seed)
    node = SceneNode()
    node.lighting = Lighting.sun_and_sky(float(rng.uniform(25, 60)), float(rng.uniform(0, 360)), sky=0.3)
    extent = 16.0
    quads = max(4, int(math.sqrt(max(8.0, 0.5 * triangles / 2))))
    terrain = heightfield_mesh(rng, quads, extent, 1.0)
    identity = node.add_transform(Transform())
    terrain_mat = node.add_material(Material(
        base_color=(0.5, 0.5, 0.5), base_color_texture=_noise_texture(rng, 64, 48),
        normal_texture=_bump_normal_map(rng, 37, 23, 2.0)))
    node.add_instance(node.add_mesh(terrain), terrain_mat, identity)

    checker = _checker(16, 16, 4, (200, 180, 60, 255), (40, 90, 160, 255))
    bumps = _bump_normal_map(rng, 8, 8, 1.5)
    box_mats = [
        node.add_material(Material(base_color_texture=checker)),
        node.add_material(Material(base_color_texture=checker, normal_texture=bumps)),
        node.add_material(Material(base_color=(0.6, 0.3, 0.2))),
        node.add_material(Material(base_color_texture=np.array([[[230, 230, 230, 255]]], dtype=np.uint8))),
    ]
    holes = _checker(32, 32, 8, (60, 160, 50, 255), (60, 160, 50, 0))
    leaf_alpha = np.clip(rng.normal(0.55, 0.35, size=(24, 40)) * 255, 0, 255).astype(np.uint8)
    card_mats = [
        node.add_material(Material(base_color_texture=holes, alpha_mode=ALPHA_MODE_MASK, alpha_cutoff=0.5,
                                   double_sided=True)),
        node.add_material(Material(base_color_texture=_noise_texture(rng, 40, 24, leaf_alpha),
                                   alpha_mode=ALPHA_MODE_MASK, alpha_cutoff=0.3, double_sided=True)),
        node.add_material(Material(base_color_texture=_noise_texture(rng, 13, 29, np.tile(leaf_alpha, (2, 1))[:29, :13]),
                                   alpha_mode=ALPHA_MODE_BLEND, double_sided=True)),
        node.add_material(Material(base_color=(0.8, 0.8, 0.9), alpha_mode=ALPHA_MODE_BLEND, double_sided=True)),
        node.add_material(Material(base_color=(0.9, 0.1, 0.1), alpha_mode=ALPHA_MODE_MASK, alpha_cutoff=1.5)),
    ]
    glow = _noise_texture(rng, 9, 5)
    sphere_mats = [
        node.add_material(Material(base_color=(0.7, 0.7, 0.7), emissive=(3.0, 2.0, 1.0), emissive_texture=glow)),
        node.add_material(Material(base_color_texture=_noise_texture(rng, 20, 20))),
    ]
    n_boxes = max(2, int(0.2 * triangles / 12))
    n_cards = max(5, int(0.1 * triangles / 2))
    n_spheres = max(2, int(0.2 * triangles / 320))
    box = node.add_mesh(box_mesh())
    for k in range(n_boxes):
        x, z = rng.uniform(-0.45 * extent, 0.45 * extent, size=2)
        sx, sz = rng.uniform(0.2, 0.8, size=2)
        h = float(min(rng.lognormal(0.0, 0.6), 4.0))
        y = _terrain_height(terrain, quads, extent, x, z) - 0.1
        node.add_instance(box, box_mats[k % len(box_mats)],
                          node.add_transform(Transform((x, y, z), (0, 0, 0), (sx, h, sz))))
    cards = [node.add_mesh(_card_mesh(r)) for r in (1.0, 2.5)]
    for k in range(n_cards):
        x, z = rng.uniform(-0.45 * extent, 0.45 * extent, size=2)
        w, h = rng.uniform(0.5, 2.0, size=2)
        y = _terrain_height(terrain, quads, extent, x, z) + 0.5 * h + rng.uniform(0.0, 0.5)
        rot = (float(rng.uniform(-20, 20)), float(rng.uniform(0, 360)), 0.0)
        node.add_instance(cards[k % 2], card_mats[k % len(card_mats)],
                          node.add_transform(Transform((x, y, z), rot, (w, h, 1.0))))
    sphere = node.add_mesh(icosphere_mesh(2))
    for k in range(n_spheres):
        r = rng.uniform(0.2, 0.8)
        x, z = rng.uniform(-0.45 * extent, 0.45 * extent, size=2)
        y = _terrain_height(terrain, quads, extent, x, z) + r + rng.uniform(0.0, 1.5)
        node.add_instance(sphere, sphere_mats[k % 2], node.add_transform(Transform((x, y, z), (0, 0, 0), (r, r, r))))
    node.camera.position = np.array([0.0, 6.0, 0.6 * extent])
    node.camera.target = np.array([0.0, 0.0, 0.0])
    node.camera.recompute()
    return node
