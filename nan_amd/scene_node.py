"""Scene graph container (scene_node.py:43-72 of the reference) plus the
lighting environment the texture-space tracer consumes.

A scene is ``meshes``, ``materials``, ``transforms`` and ``instances``
given as ``(mesh_id, material_id, transform_id)`` tuples.  The reference
reads its sun and sky from ``Scene`` (scene.py:136-147, scene.slang:7-17);
here they live on the node as a :class:`Lighting` record so that
procedural scenes carry their own environment.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .camera import Camera
from .material import Material
from .mesh import Mesh
from .transform import Transform


@dataclass
class Lighting:
    """Directional sun plus a constant sky or the physical atmosphere.

    ``sun_direction`` points toward the sun.  Sun next-event estimation
    returns ``visibility * color * intensity * transmittance * n.l``
    (scene.slang:940-995); ``sky_radiance`` replaces the sky-view LUT
    lookup on ray misses (scene.slang:469-518).
    """

    sun_direction: tuple = (0.0, 0.5, 0.866)
    sun_cos_half_angle: float = math.cos(math.radians(0.5357 / 2.0))
    sun_color: tuple = (1.0, 1.0, 1.0)
    sun_intensity: float = math.pi
    sun_transmittance: tuple = (1.0, 1.0, 1.0)
    sky_radiance: tuple = (0.0, 0.0, 0.0)
    # physical sky: misses read the sky-view table and the sun's transmittance
    # comes from the transmittance table at the view point (scene.slang:469-518,
    # 997-1009); sun_transmittance and sky_radiance are then not used
    atmosphere: bool = False

    RECORD_FLOATS = 16

    @classmethod
    def reference_default(cls) -> "Lighting":
        """The reference's default environment (scene.py:136-147): sun toward
        (0, 0.5, 0.866), intensity pi, white, 0.5357 degree disk, and the
        atmosphere."""
        return cls(atmosphere=True)

    @classmethod
    def black(cls) -> "Lighting":
        """No sun and no sky: emitters are the only light."""
        return cls(sun_intensity=0.0, sky_radiance=(0.0, 0.0, 0.0))

    @classmethod
    def sun_and_sky(cls, elevation_deg: float, azimuth_deg: float = 0.0,
                    sky: float = 0.3, intensity: float = math.pi) -> "Lighting":
        el = math.radians(elevation_deg)
        az = math.radians(azimuth_deg)
        d = (math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az))
        return cls(sun_direction=d, sun_intensity=intensity, sky_radiance=(sky, sky, sky))

    def pack(self) -> np.ndarray:
        """``float32[16]``: dir.xyz, cos_half, illuminance.rgb (color *
        intensity), pad, transmittance.rgb, pad, sky.rgb, pad."""
        rec = np.zeros(self.RECORD_FLOATS, dtype=np.float32)
        d = np.asarray(self.sun_direction, dtype=np.float64)
        n = np.linalg.norm(d)
        rec[0:3] = d / n if n > 0 else d
        rec[3] = self.sun_cos_half_angle
        rec[4:7] = np.asarray(self.sun_color, dtype=np.float64) * self.sun_intensity
        rec[8:11] = self.sun_transmittance
        rec[12:15] = self.sky_radiance
        return rec


class SceneNode:
    """Meshes, materials, transforms and instances (scene_node.py:43-72)."""

    def __init__(self):
        self.camera = Camera()
        self.materials: list[Material] = []
        self.meshes: list[Mesh] = []
        self.transforms: list[Transform] = []
        self.instances: list[tuple[int, int, int]] = []
        self.lighting = Lighting()
        self.asset_path: str | None = None
        # Russian roulette on the texture-space paths (an extension: the
        # reference's paths always run to max_bounces)
        self.russian_roulette = False

    def add_material(self, material: Material) -> int:
        self.materials.append(material)
        return len(self.materials) - 1

    def add_mesh(self, mesh: Mesh) -> int:
        self.meshes.append(mesh)
        return len(self.meshes) - 1

    def add_transform(self, transform: Transform) -> int:
        self.transforms.append(transform)
        return len(self.transforms) - 1

    def add_instance(self, mesh_id: int, material_id: int, transform_id: int) -> int:
        if not (0 <= mesh_id < len(self.meshes)):
            raise ValueError("mesh_id out of range")
        if not (0 <= material_id < len(self.materials)):
            raise ValueError("material_id out of range")
        if not (0 <= transform_id < len(self.transforms)):
            raise ValueError("transform_id out of range")
        self.instances.append((int(mesh_id), int(material_id), int(transform_id)))
        return len(self.instances) - 1

    @property
    def triangle_count(self) -> int:
        return sum(self.meshes[m].triangle_count for m, _, _ in self.instances)

    def add_scene_node(self, node: "SceneNode", transform: Transform) -> list[int]:
        """Copies ``node``'s meshes and materials in and instances every one of
        its instances under the one new ``transform`` (scene_node.py:682-718;
        the node's own instance transforms are not composed, as there)."""
        tid = self.add_transform(transform)
        mesh_base = len(self.meshes)
        mat_base = len(self.materials)
        self.meshes.extend(node.meshes)
        self.materials.extend(node.materials)
        return [self.add_instance(mesh_base + m, mat_base + k, tid) for m, k, _ in node.instances]

    @classmethod
    def load_asset(cls, path: str, default_material: Material | None = None, scale: float = 1.0,
                   auto_center: bool = False) -> "SceneNode":
        """A scene from an OBJ, glTF/GLB, STL or PLY file
        (scene_node.py:476-647); see :mod:`nan_amd.asset_import`."""
        from . import asset_import

        return asset_import.load_asset(cls, path, default_material, scale, auto_center)

    @classmethod
    def load_json(cls, path: str, axis_conversion: str | None = None) -> "SceneNode":
        """A scene composed of asset files (called at app.py:151); see
        :func:`nan_amd.asset_import.load_json` for the format."""
        from . import asset_import

        return asset_import.load_json(cls, path, axis_conversion)

    @classmethod
    def demo(cls) -> "SceneNode":
        """The reference's demo Cornell box structure (scene_node.py:721-830):
        5 double-sided walls, a ceiling light, a tall and a short block,
        8 instances and 36 triangles.  The block rotation angles are passed
        as radians, as ``spy.math.matrix_from_rotation_xyz`` receives them."""
        node = cls()
        node.camera.target = np.array([0.0, 1.0, 0.0])
        node.camera.position = np.array([0.0, 1.0, 2.95])
        node.camera.fov = 60.0
        node.camera.recompute()
        white = node.add_material(Material(base_color=(0.73, 0.73, 0.73), double_sided=True))
        red = node.add_material(Material(base_color=(0.65, 0.05, 0.05), double_sided=True))
        green = node.add_material(Material(base_color=(0.12, 0.45, 0.15), double_sided=True))
        white_box = node.add_material(Material(base_color=(0.73, 0.73, 0.73)))
        light = node.add_material(Material(base_color=(1.0, 1.0, 1.0), emissive=(15.0, 15.0, 15.0)))
        quad = node.add_mesh(Mesh.create_quad((2, 2)))
        back = node.add_mesh(Mesh.create_quad_xy((2, 2), face_positive_z=True))
        left = node.add_mesh(Mesh.create_quad_yz((2, 2), face_positive_x=True))
        right = node.add_mesh(Mesh.create_quad_yz((2, 2), face_positive_x=False))
        node.add_instance(quad, white, node.add_transform(Transform((0, 0, 0))))
        node.add_instance(quad, white, node.add_transform(Transform((0, 2, 0), (math.pi, 0, 0))))
        node.add_instance(back, white, node.add_transform(Transform((0, 1, -1))))
        node.add_instance(left, red, node.add_transform(Transform((-1, 1, 0))))
        node.add_instance(right, green, node.add_transform(Transform((1, 1, 0))))
        light_mesh = node.add_mesh(Mesh.create_quad((0.4, 0.4)))
        node.add_instance(light_mesh, light, node.add_transform(Transform((0, 1.99, 0), (math.pi, 0, 0))))
        tall = node.add_mesh(Mesh.create_cube((0.5, 1.2, 0.5)))
        node.add_instance(tall, white_box, node.add_transform(Transform((-0.35, 0.6, -0.3), (0, 18, 0))))
        short = node.add_mesh(Mesh.create_cube((0.5, 0.6, 0.5)))
        node.add_instance(short, white_box, node.add_transform(Transform((0.35, 0.3, 0.3), (0, -18, 0))))
        return node
