"""Writers for the asset-import tests: a SceneNode written as glTF/GLB, OBJ,
STL and PLY with numpy, so the importer can be checked against scenes whose
contents are known exactly.  Test helpers only; nothing in the package
imports this module."""

from __future__ import annotations

import base64
import json
import os
import struct

import numpy as np


def _col_major(m: np.ndarray) -> list:
    return [float(x) for x in np.asarray(m, dtype=np.float32).T.reshape(-1)]


def gltf_document(node, *, mesh_names=None, node_names=None):
    """(doc, blob): one glTF mesh per SceneNode mesh (positions, normals,
    uv as the scene holds it, uint32 indices), one material per
    mesh taken from the first instance that uses it, one node per instance
    carrying the instance's float32 matrix."""
    mesh_mat = {}
    for m, k, _ in node.instances:
        mesh_mat.setdefault(m, k)
    blob = bytearray()
    views, accessors, meshes, materials = [], [], [], []

    def add(arr: np.ndarray, ctype: int, typ: str, target: int) -> int:
        data = np.ascontiguousarray(arr).tobytes()
        while len(blob) % 4:
            blob.append(0)
        views.append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(data), "target": target})
        blob.extend(data)
        acc = {"bufferView": len(views) - 1, "componentType": ctype, "count": int(arr.shape[0]), "type": typ}
        if typ == "VEC3" and ctype == 5126:
            acc["min"] = [float(x) for x in arr.min(axis=0)]
            acc["max"] = [float(x) for x in arr.max(axis=0)]
        accessors.append(acc)
        return len(accessors) - 1

    for mi, mesh in enumerate(node.meshes):
        v = mesh.vertices
        uv = v[:, 6:8].copy()
        prim = {
            "attributes": {
                "POSITION": add(v[:, 0:3].copy(), 5126, "VEC3", 34962),
                "NORMAL": add(v[:, 3:6].copy(), 5126, "VEC3", 34962),
                "TEXCOORD_0": add(uv, 5126, "VEC2", 34962),
            },
            "indices": add(mesh.indices.reshape(-1, 1), 5125, "SCALAR", 34963),
            "mode": 4,
        }
        if mi in mesh_mat:
            mat = node.materials[mesh_mat[mi]]
            d = {"name": f"mat{mi}", "pbrMetallicRoughness": {"baseColorFactor": [*mat.base_color, 1.0]},
                 "doubleSided": mat.double_sided}
            if any(mat.emissive):
                peak = max(mat.emissive)
                d["emissiveFactor"] = [e / peak for e in mat.emissive]
                d["extensions"] = {"KHR_materials_emissive_strength": {"emissiveStrength": peak}}
            materials.append(d)
            prim["material"] = len(materials) - 1
        meshes.append({"name": mesh_names[mi] if mesh_names else f"m{mi}", "primitives": [prim]})
    nodes = []
    for i, (m, _, t) in enumerate(node.instances):
        nodes.append({"name": node_names[i] if node_names else f"n{i:02d}", "mesh": m,
                      "matrix": _col_major(node.transforms[t].matrix)})
    doc = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": list(range(len(nodes)))}],
           "nodes": nodes, "meshes": meshes, "materials": materials, "accessors": accessors,
           "bufferViews": views, "buffers": [{"byteLength": len(blob)}],
           "extensionsUsed": ["KHR_materials_emissive_strength"]}
    return doc, bytes(blob)


def write_gltf(path: str, doc: dict, blob: bytes, *, embed: bool = True) -> None:
    doc = json.loads(json.dumps(doc))
    if embed:
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
    else:
        name = os.path.splitext(os.path.basename(path))[0] + ".bin"
        doc["buffers"][0]["uri"] = name
        with open(os.path.join(os.path.dirname(path), name), "wb") as fh:
            fh.write(blob)
    with open(path, "w") as fh:
        json.dump(doc, fh)


def write_glb(path: str, doc: dict, blob: bytes) -> None:
    js = json.dumps(doc).encode()
    js += b" " * (-len(js) % 4)
    bb = bytes(blob) + b"\0" * (-len(blob) % 4)
    total = 12 + 8 + len(js) + 8 + len(bb)
    with open(path, "wb") as fh:
        fh.write(struct.pack("<4sII", b"glTF", 2, total))
        fh.write(struct.pack("<II", len(js), 0x4E4F534A) + js)
        fh.write(struct.pack("<II", len(bb), 0x004E4942) + bb)


def world_triangles(node):
    """Per instance: float32 world positions [T, 3, 3] (float64 product,
    rounded once) and the material index."""
    out = []
    for m, k, t in node.instances:
        mesh = node.meshes[m]
        p = mesh.vertices[:, 0:3].astype(np.float64)
        M = node.transforms[t].matrix.astype(np.float64)
        w = (p @ M[:3, :3].T + M[:3, 3]).astype(np.float32)
        out.append((w[mesh.indices.astype(np.int64)], k))
    return out


def write_obj(path: str, node, *, material_names=None) -> None:
    """World-space OBJ + MTL: one ``usemtl`` group per material, positions
    written with enough digits to read back the same float32."""
    names = material_names or [f"mat{k}" for k in range(len(node.materials))]
    mtl = os.path.splitext(path)[0] + ".mtl"
    with open(mtl, "w") as fh:
        for k, mat in enumerate(node.materials):
            fh.write(f"newmtl {names[k]}\nKd {mat.base_color[0]!r} {mat.base_color[1]!r} {mat.base_color[2]!r}\n")
            if any(mat.emissive):
                fh.write(f"Ke {mat.emissive[0]!r} {mat.emissive[1]!r} {mat.emissive[2]!r}\n")
    lines = [f"mtllib {os.path.basename(mtl)}"]
    count = 0
    for tris, k in world_triangles(node):
        lines.append(f"usemtl {names[k]}")
        for tri in tris:
            for p in tri:
                lines.append("v " + " ".join(repr(float(x)) for x in p))
            lines.append(f"f {count + 1} {count + 2} {count + 3}")
            count += 3
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def write_stl_binary(path: str, tris: np.ndarray) -> None:
    tris = np.asarray(tris, dtype=np.float32).reshape(-1, 3, 3)
    rec = np.zeros(len(tris), dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    rec["v"] = tris
    with open(path, "wb") as fh:
        fh.write(b"\0" * 80 + struct.pack("<I", len(tris)) + rec.tobytes())


def write_stl_ascii(path: str, tris: np.ndarray) -> None:
    tris = np.asarray(tris, dtype=np.float32).reshape(-1, 3, 3)
    out = ["solid test"]
    for t in tris:
        out.append("facet normal 0 0 0\nouter loop")
        out += ["vertex " + " ".join(repr(float(x)) for x in p) for p in t]
        out.append("endloop\nendfacet")
    out.append("endsolid test")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def write_ply(path: str, positions, normals, faces, *, binary: bool) -> None:
    positions = np.asarray(positions, dtype=np.float32)
    normals = np.asarray(normals, dtype=np.float32)
    head = ["ply", "format binary_little_endian 1.0" if binary else "format ascii 1.0",
            f"element vertex {len(positions)}", "property float x", "property float y", "property float z",
            "property float nx", "property float ny", "property float nz",
            f"element face {len(faces)}", "property list uchar int vertex_indices", "end_header"]
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode())
        if binary:
            fh.write(np.concatenate([positions, normals], axis=1).astype("<f4").tobytes())
            for f in faces:
                fh.write(struct.pack("<B", len(f)) + np.asarray(f, dtype="<i4").tobytes())
        else:
            for p, n in zip(positions, normals):
                fh.write((" ".join(repr(float(x)) for x in (*p, *n)) + "\n").encode())
            for f in faces:
                fh.write((f"{len(f)} " + " ".join(str(int(i)) for i in f) + "\n").encode())


def png_bytes(img: np.ndarray) -> bytes:
    """PNG encoding of an RGBA8 image (PIL)."""
    import io

    from PIL import Image

    out = io.BytesIO()
    Image.fromarray(np.ascontiguousarray(img, dtype=np.uint8), "RGBA").save(out, format="PNG")
    return out.getvalue()


def add_textures(doc: dict, blob: bytes, textures: dict, *, where: str = "bufferView", folder: str | None = None):
    """Attaches images to the glTF materials of ``doc``: ``textures`` maps
    (material index, slot) -> RGBA8 image, the slot one of
    "baseColorTexture", "emissiveTexture", "normalTexture".  Images go into
    the binary buffer (``where="bufferView"``), a data URI, or PNG files in
    ``folder`` (``where="file"``).  Returns the new blob."""
    blob = bytearray(blob)
    doc.setdefault("images", [])
    doc.setdefault("textures", [])
    for (mat, slot), img in textures.items():
        data = png_bytes(img)
        k = len(doc["images"])
        if where == "bufferView":
            while len(blob) % 4:
                blob.append(0)
            doc["bufferViews"].append({"buffer": 0, "byteOffset": len(blob), "byteLength": len(data)})
            blob.extend(data)
            doc["images"].append({"bufferView": len(doc["bufferViews"]) - 1, "mimeType": "image/png"})
        elif where == "data":
            doc["images"].append({"uri": "data:image/png;base64," + base64.b64encode(data).decode()})
        else:
            name = f"tex{k}.png"
            with open(os.path.join(folder, name), "wb") as fh:
                fh.write(data)
            doc["images"].append({"uri": name})
        doc["textures"].append({"source": k})
        m = doc["materials"][mat]
        if slot == "baseColorTexture":
            m.setdefault("pbrMetallicRoughness", {})[slot] = {"index": len(doc["textures"]) - 1}
        else:
            m[slot] = {"index": len(doc["textures"]) - 1}
    doc["buffers"][0]["byteLength"] = len(blob)
    return bytes(blob)
