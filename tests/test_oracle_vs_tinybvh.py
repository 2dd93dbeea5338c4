"""The oracle's visibility against the reference's own BVH library.

oracle/_ref/libtinybvh_ref.so is compiled (``make -C oracle ref``) from the
tinybvh the reference vendors (vertex_baker/cpp/third_party/tinybvh, used by
vertex_baker/cpp/tinybvh_visibility.cpp:145-230).  Its IsOccluded and the
oracle's any-hit test must agree on every ray whose nearest FP64 hit is not
within a hair of the ray's end, and on every ray that misses by a clear
margin; the rest are ties that two FP32 implementations may decide apart.
Skipped when the reference tree was not available to build the harness.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

import oracle
from nan_amd.scene_arrays import flatten_scene
from nan_amd.scenes import batch_scene, cornell_box

LIB = Path(__file__).resolve().parents[1] / "oracle" / "_ref" / "libtinybvh_ref.so"

pytestmark = pytest.mark.skipif(not LIB.exists(), reason="tinybvh harness not built (reference tree absent)")


def _lib():
    lib = ctypes.CDLL(str(LIB))
    lib.ref_bvh_create.restype = ctypes.c_void_p
    lib.ref_bvh_create.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    lib.ref_bvh_destroy.argtypes = [ctypes.c_void_p]
    lib.ref_bvh_occluded.restype = ctypes.c_double
    lib.ref_bvh_occluded.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                     ctypes.c_int]
    return lib


def _rays(world: np.ndarray, n: int, seed: int) -> np.ndarray:
    """Rays from points on random triangles (offset along the normal) and
    from random points in the scene box, in random directions, with finite
    and unbounded extents."""
    rng = np.random.default_rng(seed)
    tri = world.reshape(-1, 3, 3).astype(np.float64)
    lo, hi = tri.reshape(-1, 3).min(0), tri.reshape(-1, 3).max(0)
    k = rng.integers(0, len(tri), size=n)
    u, v = rng.random(n), rng.random(n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    t = tri[k]
    p = t[:, 0] + u[:, None] * (t[:, 1] - t[:, 0]) + v[:, None] * (t[:, 2] - t[:, 0])
    nrm = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    side = np.sign(np.sum(d * nrm, axis=1))[:, None]
    o = p + side * nrm * 1e-3 * max(1.0, float(np.abs(hi - lo).max()))
    half = n // 2
    o[half:] = rng.uniform(lo, hi, size=(n - half, 3))
    tmax = np.where(rng.random(n) < 0.5, 1e30, rng.uniform(0.05, 1.0, n) * float(np.linalg.norm(hi - lo)))
    rays = np.zeros((n, 8), dtype=np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays[:, 7] = tmax
    return rays


@pytest.mark.parametrize("name,node", [("cornell", cornell_box()), ("mixed 20k", batch_scene(7, 20_000))])
def test_oracle_occlusion_matches_tinybvh(name, node):
    lib = _lib()
    orc = oracle.OracleScene(flatten_scene(node))
    world = np.ascontiguousarray(orc.world_triangles())
    handle = lib.ref_bvh_create(world.ctypes.data, len(world))
    try:
        rays = _rays(world, 40_000, 11)
        ref_occ = np.zeros(len(rays), dtype=np.uint32)
        lib.ref_bvh_occluded(handle, rays.ctypes.data, len(rays), ref_occ.ctypes.data, 4)
    finally:
        lib.ref_bvh_destroy(handle)
    _, _, _, _, occ = orc.intersect(rays)
    # FP64 nearest hit along the unbounded ray
    unbounded = rays.astype(np.float64)
    unbounded[:, 7] = np.inf
    t64, _, inst64, _, _, edge = orc.intersect_fp64(unbounded)
    hit64 = inst64 != 0xFFFFFFFF
    # clear cases: the nearest hit is off the triangle's edges and well inside
    # the ray's extent, or there is no hit within 1.0001 x the extent
    tmax = rays[:, 7].astype(np.float64)
    clear_hit = hit64 & (t64 < tmax * (1 - 1e-4)) & (edge > 1e-6)
    clear_miss = ~hit64 | (t64 > tmax * (1 + 1e-4))
    clear = clear_hit | clear_miss
    assert clear.mean() > 0.99
    disagree = (occ != ref_occ) & clear
    assert disagree.sum() == 0, f"{name}: {disagree.sum()} clear rays decided differently"
    assert 0.05 < occ[clear].mean() < 0.95  # both outcomes are exercised
