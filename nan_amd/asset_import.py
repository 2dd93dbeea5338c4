"""Asset import behind ``SceneNode.load_asset`` (scene_node.py:476-647 of the
reference) and the scene description ``SceneNode.load_json`` that the
reference's application calls (app.py:151) but does not define.

The reference reads files through trimesh (with pygltflib for texture
paths); neither is a dependency here, so the formats are read directly with
numpy: Wavefront OBJ with MTL materials, glTF 2.0 (``.gltf`` with external
or data-URI buffers, and binary ``.glb``), binary and ASCII STL, and
ASCII or binary PLY.

What is kept from the reference:

* A file that holds several geometries (glTF, an OBJ with several
  materials) keeps its instancing (scene_node.py:536-606): one Mesh and one
  Material per referenced geometry, in sorted geometry-name order, and one
  instance per geometry node, in sorted node-name order, with the transform
  ``G @ world(node)`` where ``G = diag(scale) @ translate(-centre)`` (the
  centring only with ``auto_center``); the camera frames the transformed
  scene bounds.
* A file with one geometry (STL, PLY, a one-material OBJ) is centred on
  its bounds and scaled in the vertices themselves, with an identity
  transform (scene_node.py:618-640, 351-361), and the camera frames the
  moved vertices (scene_node.py:364-389).
* Camera framing: target = bounds centre, position = centre +
  (0, 0.3 d, 1.5 d), d = |max - min|.
* glTF texture coordinates reach the shader as the file stores them:
  trimesh's glTF reader turns v into 1 - v (its origin is bottom-left) and
  scene_node.py:90-95 turns it back for glTF files; OBJ, PLY and STL
  coordinates pass unchanged.
* Materials: base colour, emission (values above one kept), roughness,
  metallic, alpha mode and cutoff, double-sidedness and textures
  (scene_node.py:120-330); a geometry without a material takes
  ``default_material`` when one is given.  Textures are file paths, or
  the encoded bytes of an image a GLB or data URI embeds, decoded when the
  scene is flattened (``textures.load_image``).

Deliberate differences, all in what trimesh would have done:

* OBJ ``Kd`` is read at full precision (trimesh keeps it as 8-bit colour)
  and ``Ke`` is honoured (trimesh's OBJ material drops it), so an emitter
  written to OBJ stays an emitter.
* A texture whose file is missing stays unset (the reference substitutes
  a default texture, texture_manager.py:129-140, which for base colour is
  the factor's look).
* Missing vertex normals are the angle-weighted mean of the adjacent face
  normals; files without shared vertices (STL) have equal positions merged
  first, as trimesh's processing does.

Results against trimesh itself are unpinned (trimesh is absent); the
importer is pinned by known answers: the reference's demo scene written
out in each format loads back to the same meshes, instances and Mesh Colors
layout (216 surface texels / 276 slots / 60 double-sided at one texel per
unit, test_mesh_colors.py:107-136), and by hand-made files whose contents
are known.
"""

from __future__ import annotations

import base64
import json
import math
import os
import struct
from dataclasses import dataclass, field

import numpy as np

from .material import ALPHA_MODE_BLEND, ALPHA_MODE_MASK, ALPHA_MODE_OPAQUE, Material
from .mesh import Mesh
from .transform import Transform

SUPPORTED_EXTENSIONS = (".obj", ".gltf", ".glb", ".stl", ".ply")


@dataclass
class _Geometry:
    """One triangle list with its material description, before conversion."""

    name: str
    positions: np.ndarray  # float32[V, 3]
    faces: np.ndarray  # int64[T, 3]
    normals: np.ndarray | None = None  # float32[V, 3]
    uvs: np.ndarray | None = None  # float32[V, 2]
    colors: np.ndarray | None = None  # float32[V, >=3] in [0, 1]
    material: dict | None = None  # material description, see _make_material


@dataclass
class _Loaded:
    """A file's geometries and the nodes that place them."""

    geometries: dict = field(default_factory=dict)  # name -> _Geometry
    nodes: list = field(default_factory=list)  # (node name, float64[4,4] world, geometry name)
    single: bool = False  # one geometry, no scene graph


# ---------------------------------------------------------------------------
# Normals
# ---------------------------------------------------------------------------
def angle_weighted_normals(positions: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """Unit vertex normals: each face's unit normal weighted by the angle the
    face makes at the vertex, summed and normalised (float64, rounded once).
    Vertices on no face, or on degenerate faces only, get (0, 0, 1)."""
    p = positions.astype(np.float64)
    out = np.zeros_like(p)
    if len(faces):
        f = faces.astype(np.int64)
        a, b, c = p[f[:, 0]], p[f[:, 1]], p[f[:, 2]]
        n = np.cross(b - a, c - a)
        ln = np.linalg.norm(n, axis=1)
        ok = ln > 0.0
        n[ok] /= ln[ok, None]
        n[~ok] = 0.0

        def angle(u, v):
            lu = np.linalg.norm(u, axis=1)
            lv = np.linalg.norm(v, axis=1)
            d = np.where((lu > 0) & (lv > 0), (u * v).sum(axis=1) / np.maximum(lu * lv, 1e-300), 1.0)
            return np.arccos(np.clip(d, -1.0, 1.0))

        for k, (o, x, y) in enumerate(((a, b, c), (b, c, a), (c, a, b))):
            w = angle(x - o, y - o)
            np.add.at(out, f[:, k], n * w[:, None])
    ln = np.linalg.norm(out, axis=1)
    ok = ln > 0.0
    out[ok] /= ln[ok, None]
    out[~ok] = (0.0, 0.0, 1.0)
    return out.astype(np.float32)


def _merge_equal_positions(positions: np.ndarray, faces: np.ndarray):
    """Vertices with bit-equal positions merged, in first-appearance order."""
    key = np.ascontiguousarray(positions.astype(np.float32)).view(np.dtype((np.void, 12))).ravel()
    _, first, inverse = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return positions[first[order]], rank[inverse.ravel()][faces]


# ---------------------------------------------------------------------------
# Wavefront OBJ + MTL
# ---------------------------------------------------------------------------
def _parse_mtl(path: str) -> dict:
    mats: dict = {}
    cur = None
    base = os.path.dirname(path)
    with open(path, "r", errors="replace") as fh:
        for raw in fh:
            tok = raw.split("#", 1)[0].split()
            if not tok:
                continue
            key, args = tok[0], tok[1:]
            if key == "newmtl":
                name = " ".join(args)
                cur = mats.setdefault(name, {"name": name})
            elif cur is None:
                continue
            elif key in ("Kd", "Ke", "Ks") and len(args) >= 3:
                cur[key] = tuple(float(v) for v in args[:3])
            elif key in ("d", "Tr", "Pr", "Pm") and args:
                cur[key] = float(args[0])
            elif key in ("map_Kd", "map_Ke", "map_Bump", "bump", "norm", "map_Pr", "map_Pm", "map_Ks") and args:
                cur[key] = os.path.normpath(os.path.join(base, args[-1]))
    out = {}
    for name, m in mats.items():
        alpha = m.get("d", 1.0 - m.get("Tr", 0.0))
        desc = {
            "name": name,
            "base_color": m.get("Kd", (0.8, 0.8, 0.8)),
            "emissive": m.get("Ke", (0.0, 0.0, 0.0)),
            # trimesh's OBJ materials carry no alpha mode, so the reference
            # traces them opaque (scene_node.py:233-241 reads alphaMode only)
            "alpha_mode": ALPHA_MODE_OPAQUE,
            "dissolve": alpha,
            "base_color_texture": m.get("map_Kd"),
            "normal_texture": m.get("map_Bump") or m.get("bump") or m.get("norm"),
            "emissive_texture": m.get("map_Ke"),
            "roughness_texture": m.get("map_Pr"),
            "metallic_texture": m.get("map_Pm"),
            "specular_color_texture": m.get("map_Ks"),
        }
        if "Ks" in m:
            desc["specular_color"] = m["Ks"]
        if "Pr" in m:
            desc["roughness"] = m["Pr"]
        if "Pm" in m:
            desc["metallic"] = m["Pm"]
        out[name] = desc
    return out


def _obj_index(tok: str, count: int) -> int:
    i = int(tok)
    if i < 0:
        i += count
    else:
        i -= 1
    if not 0 <= i < count:
        raise ValueError(f"OBJ index {tok} out of range (have {count})")
    return i


def load_obj(path: str) -> _Loaded:
    """OBJ geometry grouped by material (``usemtl``), the grouping trimesh
    applies.  Polygons are split into a fan around their first corner; a
    vertex is one distinct (position, uv, normal) index triple per group."""
    v: list = []
    vt: list = []
    vn: list = []
    groups: dict = {}  # material name -> list of corner triples per polygon
    order: list = []
    materials: dict = {}
    cur = None
    base = os.path.dirname(path)
    with open(path, "r", errors="replace") as fh:
        pending = ""
        for raw in fh:
            line = pending + raw.split("#", 1)[0].rstrip("\r\n")
            if line.endswith("\\"):
                pending = line[:-1] + " "
                continue
            pending = ""
            tok = line.split()
            if not tok:
                continue
            key = tok[0]
            if key == "v":
                v.append((float(tok[1]), float(tok[2]), float(tok[3])))
            elif key == "vt":
                vt.append((float(tok[1]), float(tok[2]) if len(tok) > 2 else 0.0))
            elif key == "vn":
                vn.append((float(tok[1]), float(tok[2]), float(tok[3])))
            elif key == "f":
                corners = []
                for c in tok[1:]:
                    parts = c.split("/")
                    pi = _obj_index(parts[0], len(v))
                    ti = _obj_index(parts[1], len(vt)) if len(parts) > 1 and parts[1] else -1
                    ni = _obj_index(parts[2], len(vn)) if len(parts) > 2 and parts[2] else -1
                    corners.append((pi, ti, ni))
                if len(corners) < 3:
                    continue
                name = cur if cur is not None else ""
                if name not in groups:
                    groups[name] = []
                    order.append(name)
                groups[name].append(corners)
            elif key == "usemtl":
                cur = " ".join(tok[1:])
            elif key == "mtllib":
                for lib in tok[1:]:
                    p = os.path.join(base, lib)
                    if os.path.exists(p):
                        materials.update(_parse_mtl(p))
    pos = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    uv = np.asarray(vt, dtype=np.float64).reshape(-1, 2)
    nrm = np.asarray(vn, dtype=np.float64).reshape(-1, 3)
    out = _Loaded()
    for gname in order:
        polys = groups[gname]
        remap: dict = {}
        keys: list = []
        tris: list = []
        for corners in polys:
            ids = []
            for c in corners:
                k = remap.get(c)
                if k is None:
                    k = remap[c] = len(keys)
                    keys.append(c)
                ids.append(k)
            for j in range(1, len(ids) - 1):
                tris.append((ids[0], ids[j], ids[j + 1]))
        kk = np.asarray(keys, dtype=np.int64).reshape(-1, 3)
        faces = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
        has_uv = bool(len(kk)) and (kk[:, 1] >= 0).all()
        has_n = bool(len(kk)) and (kk[:, 2] >= 0).all()
        name = gname or "geometry"
        out.geometries[name] = _Geometry(
            name=name,
            positions=pos[kk[:, 0]].astype(np.float32),
            faces=faces,
            normals=nrm[kk[:, 2]].astype(np.float32) if has_n else None,
            uvs=uv[kk[:, 1]].astype(np.float32) if has_uv else None,
            material=materials.get(gname),
        )
        out.nodes.append((name, np.eye(4), name))
    out.single = len(out.geometries) == 1
    return out


# ---------------------------------------------------------------------------
# STL
# ---------------------------------------------------------------------------
def load_stl(path: str) -> _Loaded:
    """Binary or ASCII STL; equal positions are merged into shared vertices."""
    with open(path, "rb") as fh:
        data = fh.read()
    tri = None
    if len(data) >= 84:
        (n,) = struct.unpack_from("<I", data, 80)
        if len(data) == 84 + 50 * n:
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            tri = rec["v"].reshape(-1, 3).astype(np.float32)
    if tri is None:
        text = data.decode("ascii", errors="replace")
        if not text.lstrip().lower().startswith("solid"):
            raise ValueError(f"not an STL file: {path}")
        vals = [tuple(float(x) for x in line.split()[1:4]) for line in text.splitlines()
                if line.strip().lower().startswith("vertex")]
        if len(vals) % 3:
            raise ValueError(f"ASCII STL with {len(vals)} vertices is not a triangle list")
        tri = np.asarray(vals, dtype=np.float32).reshape(-1, 3)
    faces = np.arange(len(tri), dtype=np.int64).reshape(-1, 3)
    pos, faces = _merge_equal_positions(tri, faces)
    out = _Loaded(single=True)
    name = os.path.splitext(os.path.basename(path))[0] or "geometry"
    out.geometries[name] = _Geometry(name=name, positions=pos.astype(np.float32), faces=faces)
    out.nodes.append((name, np.eye(4), name))
    return out


# ---------------------------------------------------------------------------
# PLY
# ---------------------------------------------------------------------------
_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
              "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
              "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def load_ply(path: str) -> _Loaded:
    """ASCII or binary (either byte order) PLY: ``vertex`` x/y/z with
    optional normals, texture coordinates and colours, ``face`` index lists
    (fan-split); other elements are skipped."""
    with open(path, "rb") as fh:
        data = fh.read()
    end = data.find(b"end_header")
    if not data.startswith(b"ply") or end < 0:
        raise ValueError(f"not a PLY file: {path}")
    body = data.index(b"\n", end) + 1
    header = data[:end].decode("ascii", errors="replace").splitlines()
    fmt = None
    elements: list = []
    for line in header:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property" and elements:
            if tok[1] == "list":
                elements[-1][2].append((tok[4], "list", _PLY_TYPES[tok[2]], _PLY_TYPES[tok[3]]))
            else:
                elements[-1][2].append((tok[2], _PLY_TYPES[tok[1]]))
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"unsupported PLY format {fmt!r}")
    bo = "<" if fmt == "binary_little_endian" else ">"
    verts = None
    faces_out: list = []
    if fmt == "ascii":
        lines = data[body:].decode("ascii", errors="replace").split("\n")
        li = 0
        for name, count, props in elements:
            rows = []
            for _ in range(count):
                while not lines[li].strip():
                    li += 1
                vals = lines[li].split()
                li += 1
                rows.append(vals)
            if name == "vertex":
                cols = [p[0] for p in props]
                arr = np.asarray([[float(x) for x in r[:len(cols)]] for r in rows], dtype=np.float64)
                verts = {c: arr[:, i] for i, c in enumerate(cols)}
            elif name == "face":
                for r in rows:
                    k = int(r[0])
                    faces_out.append([int(x) for x in r[1:1 + k]])
    else:
        off = body
        for name, count, props in elements:
            if all(len(p) == 2 for p in props):
                dt = np.dtype([(p[0], bo + p[1]) for p in props])
                arr = np.frombuffer(data, dtype=dt, count=count, offset=off)
                off += dt.itemsize * count
                if name == "vertex":
                    verts = {c: arr[c].astype(np.float64) for c in dt.names}
                continue
            if name != "face" or len(props) != 1:
                raise ValueError(f"unsupported PLY element {name!r} with list properties")
            _, _, ct, it = props[0]
            cdt, idt = np.dtype(bo + ct), np.dtype(bo + it)
            for _ in range(count):
                k = int(np.frombuffer(data, dtype=cdt, count=1, offset=off)[0])
                off += cdt.itemsize
                faces_out.append(np.frombuffer(data, dtype=idt, count=k, offset=off).astype(np.int64).tolist())
                off += idt.itemsize * k
    if verts is None or not all(c in verts for c in ("x", "y", "z")):
        raise ValueError(f"PLY without vertex positions: {path}")
    pos = np.stack([verts["x"], verts["y"], verts["z"]], axis=1).astype(np.float32)
    tris = [(f[0], f[j], f[j + 1]) for f in faces_out for j in range(1, len(f) - 1)]
    faces = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    if len(faces) and (faces.min() < 0 or faces.max() >= len(pos)):
        raise ValueError("PLY face index out of range")
    nrm = uv = col = None
    if all(c in verts for c in ("nx", "ny", "nz")):
        nrm = np.stack([verts["nx"], verts["ny"], verts["nz"]], axis=1).astype(np.float32)
    for a, b in (("u", "v"), ("s", "t"), ("texture_u", "texture_v")):
        if a in verts and b in verts:
            uv = np.stack([verts[a], verts[b]], axis=1).astype(np.float32)
            break
    if all(c in verts for c in ("red", "green", "blue")):
        col = np.stack([verts["red"], verts["green"], verts["blue"]], axis=1) / 255.0
    out = _Loaded(single=True)
    name = os.path.splitext(os.path.basename(path))[0] or "geometry"
    out.geometries[name] = _Geometry(name=name, positions=pos, faces=faces, normals=nrm, uvs=uv, colors=col)
    out.nodes.append((name, np.eye(4), name))
    return out


# ---------------------------------------------------------------------------
# glTF 2.0
# ---------------------------------------------------------------------------
_GLTF_COMPONENT = {5120: "i1", 5121: "u1", 5122: "<i2", 5123: "<u2", 5125: "<u4", 5126: "<f4"}
_GLTF_WIDTH = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}
_GLTF_NORMALIZE = {5120: 127.0, 5121: 255.0, 5122: 32767.0, 5123: 65535.0}


class _Gltf:
    def __init__(self, path: str):
        self.base = os.path.dirname(os.path.abspath(path))
        with open(path, "rb") as fh:
            data = fh.read()
        self.bin = None
        if data[:4] == b"glTF":
            version, length = struct.unpack_from("<II", data, 4)
            if version != 2:
                raise ValueError(f"glTF binary version {version} is not supported")
            off, doc = 12, None
            while off + 8 <= min(length, len(data)):
                clen, ctype = struct.unpack_from("<II", data, off)
                chunk = data[off + 8: off + 8 + clen]
                if ctype == 0x4E4F534A:
                    doc = json.loads(chunk.decode("utf-8"))
                elif ctype == 0x004E4942 and self.bin is None:
                    self.bin = chunk
                off += 8 + clen + (-clen % 4)
            if doc is None:
                raise ValueError("glTF binary without a JSON chunk")
            self.doc = doc
        else:
            self.doc = json.loads(data.decode("utf-8"))
        ver = str(self.doc.get("asset", {}).get("version", "2.0"))
        if not ver.startswith("2"):
            raise ValueError(f"glTF version {ver} is not supported")
        self._buffers: dict = {}

    def buffer(self, i: int) -> bytes:
        if i not in self._buffers:
            b = self.doc["buffers"][i]
            uri = b.get("uri")
            if uri is None:
                if self.bin is None:
                    raise ValueError(f"buffer {i} has no uri and the file has no binary chunk")
                raw = self.bin
            elif uri.startswith("data:"):
                raw = base64.b64decode(uri.split(",", 1)[1])
            else:
                from urllib.parse import unquote
                with open(os.path.join(self.base, unquote(uri)), "rb") as fh:
                    raw = fh.read()
            if len(raw) < b.get("byteLength", 0):
                raise ValueError(f"buffer {i} is shorter than its byteLength")
            self._buffers[i] = raw
        return self._buffers[i]

    def _view(self, view_index: int, offset: int, count: int, dt: np.dtype, width: int) -> np.ndarray:
        view = self.doc["bufferViews"][view_index]
        buf = self.buffer(view["buffer"])
        start = view.get("byteOffset", 0) + offset
        elem = dt.itemsize * width
        stride = view.get("byteStride") or elem
        if count == 0:
            return np.zeros((0, width), dtype=dt)
        need = start + stride * (count - 1) + elem
        if need > view.get("byteOffset", 0) + view["byteLength"] or need > len(buf):
            raise ValueError(f"accessor reads past bufferView {view_index}")
        if stride == elem:
            return np.frombuffer(buf, dtype=dt, count=count * width, offset=start).reshape(count, width)
        raw = np.frombuffer(buf, dtype=np.uint8, count=stride * (count - 1) + elem, offset=start)
        rows = np.lib.stride_tricks.as_strided(raw, shape=(count, elem), strides=(stride, 1))
        return np.ascontiguousarray(rows).view(dt).reshape(count, width)

    def accessor(self, i: int) -> np.ndarray:
        """float64 or int64 ``[count, width]`` with normalisation and sparse
        substitution applied."""
        a = self.doc["accessors"][i]
        ct = a["componentType"]
        dt = np.dtype(_GLTF_COMPONENT[ct])
        width = _GLTF_WIDTH[a["type"]]
        count = a["count"]
        if "bufferView" in a:
            arr = self._view(a["bufferView"], a.get("byteOffset", 0), count, dt, width).copy()
        else:
            arr = np.zeros((count, width), dtype=dt)
        sp = a.get("sparse")
        if sp:
            n = sp["count"]
            idx_dt = np.dtype(_GLTF_COMPONENT[sp["indices"]["componentType"]])
            idx = self._view(sp["indices"]["bufferView"], sp["indices"].get("byteOffset", 0), n, idx_dt, 1)[:, 0]
            vals = self._view(sp["values"]["bufferView"], sp["values"].get("byteOffset", 0), n, dt, width)
            if len(idx) and int(idx.max()) >= count:
                raise ValueError("sparse accessor index out of range")
            arr[idx.astype(np.int64)] = vals
        if ct == 5126:
            return arr.astype(np.float64)
        if a.get("normalized"):
            return np.maximum(arr.astype(np.float64) / _GLTF_NORMALIZE[ct], -1.0)
        return arr.astype(np.int64)

    def texture_path(self, info):
        """The image of a texture reference: a file path, or the encoded
        bytes of an image held in a bufferView or a data URI; None when
        there is none or its file is missing."""
        if not info or info.get("index") is None:
            return None
        tex = self.doc.get("textures", [])[info["index"]]
        src = tex.get("source")
        if src is None:
            return None
        image = self.doc.get("images", [])[src]
        if "bufferView" in image:
            view = self.doc["bufferViews"][image["bufferView"]]
            buf = self.buffer(view["buffer"])
            start = view.get("byteOffset", 0)
            return bytes(buf[start:start + view["byteLength"]])
        uri = image.get("uri")
        if not uri:
            return None
        if uri.startswith("data:"):
            return base64.b64decode(uri.split(",", 1)[1])
        from urllib.parse import unquote
        path = os.path.normpath(unquote(uri) if os.path.isabs(uri) else os.path.join(self.base, unquote(uri)))
        return path if os.path.exists(path) else None

    def material(self, i: int) -> dict:
        m = self.doc["materials"][i]
        pbr = m.get("pbrMetallicRoughness", {})
        ext = m.get("extensions", {})
        emissive = np.asarray(m.get("emissiveFactor", (0.0, 0.0, 0.0)), dtype=np.float64)
        strength = ext.get("KHR_materials_emissive_strength", {}).get("emissiveStrength")
        if strength is not None:
            emissive = emissive * float(strength)
        mode = {"OPAQUE": ALPHA_MODE_OPAQUE, "MASK": ALPHA_MODE_MASK, "BLEND": ALPHA_MODE_BLEND}[
            str(m.get("alphaMode", "OPAQUE")).upper()]
        desc = {
            "name": m.get("name"),
            "base_color": tuple(pbr.get("baseColorFactor", (1.0, 1.0, 1.0, 1.0))[:3]),
            "emissive": tuple(float(x) for x in emissive),
            "alpha_mode": mode,
            "alpha_cutoff": float(m.get("alphaCutoff", 0.5)),
            "double_sided": bool(m.get("doubleSided", False)),
            "base_color_texture": self.texture_path(pbr.get("baseColorTexture")),
            "normal_texture": self.texture_path(m.get("normalTexture")),
            "emissive_texture": self.texture_path(m.get("emissiveTexture")),
        }
        mr = self.texture_path(pbr.get("metallicRoughnessTexture"))
        desc["roughness_texture"] = desc["metallic_texture"] = mr
        if "roughnessFactor" in pbr:
            desc["roughness"] = float(pbr["roughnessFactor"])
        if "metallicFactor" in pbr:
            desc["metallic"] = float(pbr["metallicFactor"])
        sg = ext.get("KHR_materials_pbrSpecularGlossiness")
        if sg:
            if "diffuseFactor" in sg:
                desc["base_color"] = tuple(sg["diffuseFactor"][:3])
            desc["base_color_texture"] = self.texture_path(sg.get("diffuseTexture")) or desc["base_color_texture"]
        spec = ext.get("KHR_materials_specular")
        if spec:
            if spec.get("specularColorFactor") is not None:
                desc["specular_color"] = tuple(spec["specularColorFactor"][:3])
            desc["specular_color_texture"] = self.texture_path(spec.get("specularColorTexture"))
        return desc


def _node_local(node: dict) -> np.ndarray:
    if "matrix" in node:
        return np.asarray(node["matrix"], dtype=np.float64).reshape(4, 4).T  # column-major
    t = np.eye(4)
    t[:3, 3] = node.get("translation", (0.0, 0.0, 0.0))
    x, y, z, w = (float(v) for v in node.get("rotation", (0.0, 0.0, 0.0, 1.0)))
    n = math.sqrt(x * x + y * y + z * z + w * w) or 1.0
    x, y, z, w = x / n, y / n, z / n, w / n
    r = np.eye(4)
    r[:3, :3] = [[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                 [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                 [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]]
    s = np.diag([*node.get("scale", (1.0, 1.0, 1.0)), 1.0])
    return t @ r @ s


def _unique(name: str, taken: set) -> str:
    out, k = name, 1
    while out in taken:
        out = f"{name}_{k}"
        k += 1
    taken.add(out)
    return out


def _primitive_faces(mode: int, idx: np.ndarray) -> np.ndarray:
    if mode == 4:
        n = len(idx) - len(idx) % 3
        return idx[:n].reshape(-1, 3)
    if mode == 5:  # strip: alternate winding keeps every triangle facing the same way
        k = np.arange(max(len(idx) - 2, 0))
        a = np.where(k % 2 == 0, idx[k], idx[k + 1]) if len(k) else k
        b = np.where(k % 2 == 0, idx[k + 1], idx[k]) if len(k) else k
        return np.stack([a, b, idx[k + 2]], axis=1) if len(k) else np.zeros((0, 3), np.int64)
    if mode == 6:  # fan
        k = np.arange(1, max(len(idx) - 1, 1))
        return np.stack([np.full(len(k), idx[0]), idx[k], idx[k + 1]], axis=1) if len(k) else np.zeros((0, 3), np.int64)
    return np.zeros((0, 3), np.int64)  # points and lines carry no surface


def load_gltf(path: str) -> _Loaded:
    g = _Gltf(path)
    doc = g.doc
    out = _Loaded()
    taken: set = set()
    prim_names: dict = {}  # mesh index -> list of geometry names (None where skipped)
    for mi, mesh in enumerate(doc.get("meshes", [])):
        base = mesh.get("name") or f"mesh_{mi}"
        names = []
        prims = mesh.get("primitives", [])
        for pi, prim in enumerate(prims):
            attrs = prim.get("attributes", {})
            if "POSITION" not in attrs or prim.get("mode", 4) not in (4, 5, 6):
                names.append(None)
                continue
            pos = g.accessor(attrs["POSITION"])[:, :3]
            idx = g.accessor(prim["indices"])[:, 0] if "indices" in prim else np.arange(len(pos), dtype=np.int64)
            if len(idx) and (idx.min() < 0 or idx.max() >= len(pos)):
                raise ValueError(f"mesh {base!r} primitive {pi}: index out of range")
            faces = _primitive_faces(prim.get("mode", 4), idx)
            name = _unique(base if len(prims) == 1 else f"{base}_{pi}", taken)
            geo = _Geometry(name=name, positions=pos.astype(np.float32), faces=faces.astype(np.int64))
            if "NORMAL" in attrs:
                geo.normals = g.accessor(attrs["NORMAL"])[:, :3].astype(np.float32)
            if "TEXCOORD_0" in attrs:
                geo.uvs = g.accessor(attrs["TEXCOORD_0"])[:, :2].astype(np.float32)
            if "COLOR_0" in attrs:
                geo.colors = g.accessor(attrs["COLOR_0"])
            if "material" in prim:
                geo.material = g.material(prim["material"])
            out.geometries[name] = geo
            names.append(name)
        prim_names[mi] = names
    nodes = doc.get("nodes", [])
    scenes = doc.get("scenes", [])
    if scenes:
        roots = scenes[doc.get("scene", 0)].get("nodes", [])
    else:
        children = {c for n in nodes for c in n.get("children", [])}
        roots = [i for i in range(len(nodes)) if i not in children]
    node_taken: set = set()
    stack = [(r, np.eye(4), 0) for r in reversed(roots)]
    while stack:
        ni, parent, depth = stack.pop()
        if depth > len(nodes):
            raise ValueError("glTF node hierarchy has a cycle")
        node = nodes[ni]
        world = parent @ _node_local(node)
        if "mesh" in node:
            names = prim_names.get(node["mesh"], [])
            nname = node.get("name") or f"node_{ni}"
            live = [n for n in names if n is not None]
            for j, gname in enumerate(live):
                out.nodes.append((_unique(nname if len(live) == 1 else f"{nname}_{j}", node_taken), world, gname))
        for c in reversed(node.get("children", [])):
            stack.append((c, world, depth + 1))
    return out


# ---------------------------------------------------------------------------
# Conversion to SceneNode
# ---------------------------------------------------------------------------
def _make_material(desc: dict | None, colors, default: Material | None) -> Material:
    """A Material from a parsed description (scene_node.py:120-330)."""
    if desc is None:
        if default is not None:
            return default
        base = (0.8, 0.8, 0.8)
        if colors is not None and len(colors):
            base = tuple(float(x) for x in np.clip(colors[:, :3].mean(axis=0), 0.0, 1.0))
        return Material(base_color=base)
    base = np.asarray(desc.get("base_color", (0.8, 0.8, 0.8)), dtype=np.float64)

    def existing(key):
        src = desc.get(key)
        return None if isinstance(src, str) and not os.path.exists(src) else src

    mat = Material(
        base_color=tuple(float(x) for x in base),
        emissive=desc.get("emissive", (0.0, 0.0, 0.0)),
        specular_color=desc.get("specular_color", (1.0, 1.0, 1.0)),
        roughness=desc.get("roughness", 0.5),
        metallic=desc.get("metallic", 0.0),
        alpha_mode=int(desc.get("alpha_mode", ALPHA_MODE_OPAQUE)),
        alpha_cutoff=desc.get("alpha_cutoff", 0.5),
        double_sided=desc.get("double_sided", False),
        base_color_texture=existing("base_color_texture"),
        emissive_texture=existing("emissive_texture"),
        normal_texture=existing("normal_texture"),
        roughness_texture=existing("roughness_texture"),
        metallic_texture=existing("metallic_texture"),
        specular_color_texture=existing("specular_color_texture"),
    )
    mat.name = desc.get("name")
    return mat


def _make_mesh(geo: _Geometry, positions: np.ndarray) -> Mesh:
    """float32[V, 8] vertices (position, normal, uv) and uint32 indices
    (scene_node.py:73-117); uvs as the file stores them (see the module
    docstring on glTF's v)."""
    normals = geo.normals if geo.normals is not None else angle_weighted_normals(positions, geo.faces)
    verts = np.zeros((len(positions), 8), dtype=np.float32)
    verts[:, 0:3] = positions
    verts[:, 3:6] = normals
    if geo.uvs is not None:
        verts[:, 6:8] = geo.uvs
    return Mesh(verts, geo.faces.astype(np.uint32))


def _frame_camera(node, bounds_min: np.ndarray, bounds_max: np.ndarray) -> None:
    """Camera toward the bounds' centre from 0.3 d above and 1.5 d in front
    (+z), d the bounds' diagonal (scene_node.py:364-389, 608-615)."""
    center = (bounds_min + bounds_max) * 0.5
    size = float(np.linalg.norm(bounds_max - bounds_min))
    node.camera.target = center.astype(np.float32).astype(np.float64)
    node.camera.position = np.array([center[0], center[1] + size * 0.3, center[2] + size * 1.5])
    node.camera.recompute()


def load_asset(cls, path: str, default_material: Material | None = None, scale: float = 1.0,
               auto_center: bool = False):
    """``SceneNode.load_asset`` (scene_node.py:476-647)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Asset not found: {path}")
    ext = os.path.splitext(path)[1].lower()
    loaders = {".obj": load_obj, ".gltf": load_gltf, ".glb": load_gltf, ".stl": load_stl, ".ply": load_ply}
    if ext not in loaders:
        raise ValueError(f"Failed to load asset: unsupported file type {ext!r} "
                         f"(supported: {', '.join(SUPPORTED_EXTENSIONS)})")
    try:
        loaded = loaders[ext](path)
    except (OSError, KeyError, IndexError, struct.error, json.JSONDecodeError, UnicodeDecodeError) as e:
        raise ValueError(f"Failed to load asset: {e}") from e
    node = cls()
    node.asset_path = path
    scale = float(scale)

    if loaded.single:
        geo = next(iter(loaded.geometries.values()))
        if len(geo.positions) == 0 or len(geo.faces) == 0:
            raise ValueError(f"Asset contains no valid geometry: {path}")
        p = geo.positions.astype(np.float64)
        if auto_center:
            p = p - (p.min(axis=0) + p.max(axis=0)) * 0.5
        if scale != 1.0:
            p = p * scale
        pos = p.astype(np.float32)
        mesh_id = node.add_mesh(_make_mesh(geo, pos))
        mat_id = node.add_material(_make_material(geo.material, geo.colors, default_material))
        node.add_instance(mesh_id, mat_id, node.add_transform(Transform()))
        _frame_camera(node, p.min(axis=0), p.max(axis=0))
        return node

    # scene path: instancing kept
    used = {g for _, _, g in loaded.nodes}
    ids: dict = {}
    for gname in sorted(used):
        geo = loaded.geometries[gname]
        if len(geo.positions) == 0 or len(geo.faces) == 0:
            continue
        mesh_id = node.add_mesh(_make_mesh(geo, geo.positions))
        mat_id = node.add_material(_make_material(geo.material, geo.colors, default_material))
        ids[gname] = (mesh_id, mat_id)
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for _, world, gname in loaded.nodes:
        if gname not in ids:
            continue
        p = loaded.geometries[gname].positions.astype(np.float64)
        w = p @ world[:3, :3].T + world[:3, 3]
        lo = np.minimum(lo, w.min(axis=0))
        hi = np.maximum(hi, w.max(axis=0))
    if not ids:
        raise ValueError(f"Asset contains no valid geometry: {path}")
    center = (lo + hi) * 0.5
    glob = np.eye(4)
    if auto_center:
        glob[:3, 3] = -center
    if scale != 1.0:
        glob = np.diag([scale, scale, scale, 1.0]) @ glob
    for nname, world, gname in sorted(loaded.nodes, key=lambda n: n[0]):
        if gname not in ids:
            continue
        node.add_instance(ids[gname][0], ids[gname][1],
                          node.add_transform(Transform.from_matrix((glob @ world).astype(np.float32))))
    bmin = (lo - (center if auto_center else 0.0)) * scale
    bmax = (hi - (center if auto_center else 0.0)) * scale
    _frame_camera(node, np.minimum(bmin, bmax), np.maximum(bmin, bmax))
    return node


# ---------------------------------------------------------------------------
# Scene description (JSON)
# ---------------------------------------------------------------------------
def z_up_to_y_up(vec) -> list:
    """(x, y, z) -> (x, z, -y) (scene_node.py:648-662)."""
    v = list(vec)
    return [v[0], v[2], -v[1]] if len(v) == 3 else v


def rotation_z_up_to_y_up(rot) -> list:
    """Euler angles (rx, ry, rz) -> (rz, rx, -ry) (scene_node.py:664-679)."""
    r = list(rot)
    return [r[2], r[0], -r[1]] if len(r) == 3 else r


def _vec3(data: dict, key: str, default, where: str):
    v = data.get(key, default)
    if np.isscalar(v):
        v = [v, v, v]
    if not isinstance(v, (list, tuple)) or len(v) != 3:
        raise ValueError(f"{where}.{key} must be a 3-number array")
    out = [float(x) for x in v]
    if not all(math.isfinite(x) for x in out):
        raise ValueError(f"{where}.{key} contains non-finite values")
    return out


def load_json(cls, path: str, axis_conversion: str | None = None):
    """``SceneNode.load_json`` (called at app.py:151; the reference defines
    no body).  The format defined here::

        {"objects": [{"asset": "rel/path.glb", "translation": [x, y, z],
                      "rotation": [rx, ry, rz] (degrees), "scale": s or [sx, sy, sz],
                      "auto_center": false, "asset_scale": 1.0}, ...],
         "camera": {"position": [...], "target": [...], "fov": 60},
         "lighting": {"sun_direction": [...], "sun_intensity": pi,
                      "sun_color": [...], "sky_radiance": [...], "atmosphere": true}}

    Each object is loaded with :func:`load_asset` and placed by its object
    transform ``T * R * S`` (rotation X then Y then Z) composed with each of
    its own instance transforms, so an instanced asset keeps its layout (the
    reference's ``add_scene_node``, scene_node.py:682-718, gives every
    instance the one object transform instead; that method keeps that
    behaviour).
    With ``axis_conversion="z_up_to_y_up"`` the vectors and rotations of the
    description (not the asset files, which carry their own convention) are
    converted with the reference's helpers (scene_node.py:648-679).
    """
    from .scene_node import Lighting

    if axis_conversion not in (None, "z_up_to_y_up"):
        raise ValueError(f"unknown axis_conversion {axis_conversion!r}")
    conv = axis_conversion == "z_up_to_y_up"
    with open(path, "r") as fh:
        doc = json.load(fh)
    if not isinstance(doc, dict) or not isinstance(doc.get("objects"), list) or not doc["objects"]:
        raise ValueError("scene description needs a non-empty 'objects' list")
    base = os.path.dirname(os.path.abspath(path))
    node = cls()
    node.asset_path = path
    for k, obj in enumerate(doc["objects"]):
        where = f"objects[{k}]"
        if not isinstance(obj, dict) or not isinstance(obj.get("asset"), str):
            raise ValueError(f"{where}.asset must be a path")
        asset = obj["asset"] if os.path.isabs(obj["asset"]) else os.path.join(base, obj["asset"])
        child = load_asset(cls, asset, scale=float(obj.get("asset_scale", 1.0)),
                           auto_center=bool(obj.get("auto_center", False)))
        t = _vec3(obj, "translation", [0.0, 0.0, 0.0], where)
        r = _vec3(obj, "rotation", [0.0, 0.0, 0.0], where)
        s = _vec3(obj, "scale", [1.0, 1.0, 1.0], where)
        if conv:
            t, r = z_up_to_y_up(t), rotation_z_up_to_y_up(r)
        obj = Transform(t, [math.radians(a) for a in r], s).matrix.astype(np.float64)
        mesh_base, mat_base = len(node.meshes), len(node.materials)
        node.meshes.extend(child.meshes)
        node.materials.extend(child.materials)
        for m, k, ti in child.instances:
            placed = (obj @ child.transforms[ti].matrix.astype(np.float64)).astype(np.float32)
            node.add_instance(mesh_base + m, mat_base + k, node.add_transform(Transform.from_matrix(placed)))
    cam = doc.get("camera")
    if cam is not None:
        pos = _vec3(cam, "position", list(node.camera.position), "camera")
        tgt = _vec3(cam, "target", list(node.camera.target), "camera")
        if conv:
            pos, tgt = z_up_to_y_up(pos), z_up_to_y_up(tgt)
        node.camera.position = np.asarray(pos)
        node.camera.target = np.asarray(tgt)
        node.camera.fov = float(cam.get("fov", node.camera.fov))
        node.camera.recompute()
    lit = doc.get("lighting")
    if lit is not None:
        sun = _vec3(lit, "sun_direction", [0.0, 0.5, 0.866], "lighting")
        node.lighting = Lighting(
            sun_direction=tuple(z_up_to_y_up(sun) if conv else sun),
            sun_intensity=float(lit.get("sun_intensity", math.pi)),
            sun_color=tuple(_vec3(lit, "sun_color", [1.0, 1.0, 1.0], "lighting")),
            sky_radiance=tuple(_vec3(lit, "sky_radiance", [0.0, 0.0, 0.0], "lighting")),
            atmosphere=bool(lit.get("atmosphere", False)),
        )
    return node
