"""On-disk frozen cache: round trip and rejection of damaged or foreign
files (CPU)."""

from __future__ import annotations

import numpy as np
import pytest

from nan_amd.frozen_cache import FrozenCacheError, layout_digest, load_frozen_cache, save_frozen_cache
from nan_amd.mesh_colors import MeshColorsLayout
from nan_amd.scenes import batch_scene, cornell_box


@pytest.fixture(scope="module")
def layouts():
    return MeshColorsLayout.build(cornell_box()), MeshColorsLayout.build(batch_scene(3, 3000), texels_per_unit=4.0,
                                                                        min_resolution=1, max_resolution=8)


def _words(n, seed=0):
    return np.random.default_rng(seed).integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)


def test_round_trip(tmp_path, layouts):
    layout, _ = layouts
    words = _words(layout.total_payload_count)
    path = tmp_path / "cache.nmc"
    save_frozen_cache(path, layout, words)
    got, stored = load_frozen_cache(path, layout)
    np.testing.assert_array_equal(got, words)
    np.testing.assert_array_equal(stored.face_table, layout.face_table)
    np.testing.assert_array_equal(stored.instance_table, layout.instance_table)
    np.testing.assert_array_equal(stored.side_table, layout.side_table)
    assert (stored.total_surface_texels, stored.total_payload_count) == (layout.total_surface_texels,
                                                                         layout.total_payload_count)
    assert layout_digest(stored) == layout_digest(layout)
    assert path.stat().st_size == 96 + layout.face_table.nbytes + layout.instance_table.nbytes + \
        layout.side_table.nbytes + 4 * layout.total_payload_count + 32


def test_rejects_damage_and_foreign_layouts(tmp_path, layouts):
    layout, other = layouts
    path = tmp_path / "cache.nmc"
    save_frozen_cache(path, layout, _words(layout.total_payload_count, 1))
    data = bytearray(path.read_bytes())
    with pytest.raises(FrozenCacheError, match="different Mesh Colors layout"):
        load_frozen_cache(path, other)
    for offset, what in ((len(data) // 2, "checksum"), (0, "magic"), (8, "version")):
        bad = bytearray(data)
        bad[offset] ^= 0x01
        p = tmp_path / f"bad{offset}.nmc"
        p.write_bytes(bytes(bad))
        with pytest.raises(FrozenCacheError, match=what):
            load_frozen_cache(p)
    short = tmp_path / "short.nmc"
    short.write_bytes(bytes(data[:-5]))
    with pytest.raises(FrozenCacheError, match="size"):
        load_frozen_cache(short)
    with pytest.raises(ValueError):
        save_frozen_cache(tmp_path / "x.nmc", layout, _words(layout.total_payload_count - 1))
