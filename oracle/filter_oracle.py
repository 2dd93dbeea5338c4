"""CPU restatement of the Mesh Colors surface filter — TEST INFRASTRUCTURE
ONLY (tests and the smoke check load it; the product never does).

Follows mesh_colors_surface_filter.slang:17-253 and the helpers in
mesh_colors.slang: mesh_colors_get_adjacency (102-126),
mesh_colors_map_one_edge_tap (152-200), mesh_colors_normal_filter_weight
(202-214), mesh_colors_sample_irradiance (365-426).  Scalar FP32 arithmetic
(numpy float32) in the reference's operation order; acos/exp are the C
library's, so results agree with the GPU to a few ulp, not bit for bit.
A plain loop over texels: sized for test meshes and the Cornell box.
"""

from __future__ import annotations

import math

import numpy as np

F = np.float32
INVALID = 0xFFFFFFFF
OFFSETS = ((1, 0), (-1, 0), (0, 1), (0, -1), (1, -1), (-1, 1))  # surface_filter.slang:17-29


def _row_offset(i: int, r: int) -> int:
    return i * (r + 1) - (i * (i - 1)) // 2


def _address_to_ij(a: int, r: int) -> tuple[int, int]:
    i = 0
    while i < r and _row_offset(i + 1, r) <= a:
        i += 1
    return i, a - _row_offset(i, r)


def _texel(i: int, j: int, inst, face) -> int:
    """mesh_colors_address: instance texel offset + face address + lattice index."""
    return int(inst[2]) + int(face[0]) + _row_offset(i, int(face[1])) + j


def _side(front: int, inst, side, back: bool) -> int:
    return int(side[0]) + (front - int(inst[2])) if back else front


def map_one_edge_tap(ci: int, cj: int, resolution: int, adj_edge: int, flip: bool, crossed: int):
    """mesh_colors.slang:152-200; returns (valid, barycentrics float32[3])."""
    inv = F(1.0) / F(resolution)
    src = [F(ci) * inv, F(cj) * inv, F(1.0) - F(ci + cj) * inv]
    s0, s1, sop = crossed, (crossed + 1) % 3, (crossed + 2) % 3
    n0, n1, nop = adj_edge, (adj_edge + 1) % 3, (adj_edge + 2) % 3
    outside = src[sop]
    m0, m1 = (n1, n0) if flip else (n0, n1)
    nb = [F(0.0), F(0.0), F(0.0)]
    nb[m0] = src[s0] + outside
    nb[m1] = src[s1] + outside
    nb[nop] = -outside
    if any(v < F(-1e-5) for v in nb) or any(v > F(1.0 + 1e-5) for v in nb):
        return False, np.zeros(3, dtype=np.float32)
    nb = [max(v, F(0.0)) for v in nb]
    total = max((nb[0] + nb[1]) + nb[2], F(1e-6))
    return True, np.array([v / total for v in nb], dtype=np.float32)


def normal_filter_weight(center, sample, sigma: float) -> np.float32:
    """mesh_colors.slang:202-214."""
    c = F(center[0]) * F(sample[0]) + F(center[1]) * F(sample[1]) + F(center[2]) * F(sample[2])
    c = min(max(c, F(0.0)), F(1.0))
    if F(sigma) <= F(1e-6):
        return F(1.0) if c >= F(1.0 - 1e-6) else F(0.0)
    a = F(math.acos(float(c))) / F(sigma)
    return F(math.exp(float(F(-0.5) * a * a)))


def sample_irradiance(payload, bary, inst, face, side, back: bool) -> np.ndarray:
    """mesh_colors.slang:365-426 (float32)."""
    r = int(face[1])
    fr = F(r)
    sc = [min(max(F(b), F(0.0)), F(1.0)) * fr for b in bary]
    i, j = int(sc[0]), int(sc[1])
    i = min(i, r)
    j = min(j, r - i)
    w = [v - F(math.floor(float(v))) for v in sc]
    upper = abs((w[0] + w[1]) + w[2] - F(2.0)) <= F(1e-3)
    if upper:
        w = [F(1.0) - v for v in w]

    def load(ii, jj):
        return payload[_side(_texel(ii, jj, inst, face), inst, side, back), :3].astype(np.float32)

    if (w[0] + w[1]) + w[2] <= F(1e-6) or i >= r or j >= r - i:
        return load(i, j)
    if upper:
        ta, tb, tc = (i, j + 1), (i + 1, j), (i + 1, j + 1)
    else:
        ta, tb, tc = (i + 1, j), (i, j + 1), (i, j)
    return (w[0] * load(*ta) + w[1] * load(*tb)) + w[2] * load(*tc)


def surface_filter(faces, instances, sides, adjacency, normals, payload, neighbor_weight: float,
                   normal_sigma: float) -> np.ndarray:
    """One filter pass (mesh_colors_surface_filter.slang:90-253) over every
    texel of every instance; returns the destination payload (slots that no
    texel writes stay zero, as in a fresh destination buffer)."""
    faces = np.asarray(faces, dtype=np.uint32)
    instances = np.asarray(instances, dtype=np.uint32)
    sides = np.asarray(sides, dtype=np.uint32)
    adjacency = np.asarray(adjacency, dtype=np.uint32)
    normals = np.asarray(normals, dtype=np.float32)
    src = np.asarray(payload, dtype=np.float32)
    dst = np.zeros_like(src)
    nw = F(neighbor_weight)
    for k, inst in enumerate(instances):
        side = sides[k]
        double = bool(side[1] & 1)
        face_offset, face_count = int(inst[0]), int(inst[1])
        for prim in range(face_count):
            face = faces[face_offset + prim]
            r = int(face[1])
            center_normal = normals[face_offset + prim, :3]
            for a in range((r + 1) * (r + 2) // 2):
                i, j = _address_to_ij(a, r)
                front = _texel(i, j, inst, face)
                for back in ((False, True) if double else (False,)):
                    center = _side(front, inst, side, back)
                    acc = src[center, :3].astype(np.float32).copy()
                    wsum = F(1.0)
                    for di, dj in OFFSETS:
                        ci, cj = i + di, j + dj
                        if ci >= 0 and cj >= 0 and ci + cj <= r:
                            v = src[_side(_texel(ci, cj, inst, face), inst, side, back), :3]
                            acc = acc + nw * v
                            wsum = wsum + nw
                            continue
                        crossed = 1 if ci < 0 else (2 if cj < 0 else 0)
                        packed = int(adjacency[face_offset + prim, crossed])
                        if packed == INVALID:
                            continue
                        adj_face = face_offset + (packed >> 3)
                        ok, nb = map_one_edge_tap(ci, cj, r, packed & 3, bool(packed & 4), crossed)
                        if not ok:
                            continue
                        tw = nw * normal_filter_weight(center_normal, normals[adj_face, :3], normal_sigma)
                        if not tw > F(0.0):
                            continue
                        v = sample_irradiance(src, nb, inst, faces[adj_face], side, back)
                        acc = acc + tw * v
                        wsum = wsum + tw
                    dst[center, :3] = acc / max(wsum, F(1e-6))
                    dst[center, 3] = src[center, 3]
    return dst
