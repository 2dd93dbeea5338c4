"""Renderer save/freeze state machine driven by mocks, as the reference tests
it (test_mesh_colors_surface_filter.py:377-455)."""

from __future__ import annotations

from types import SimpleNamespace

from nan_amd.texture_space_path_tracing_renderer import TextureSpacePathTracingRenderer


def _renderer() -> TextureSpacePathTracingRenderer:
    r = TextureSpacePathTracingRenderer.__new__(TextureSpacePathTracingRenderer)
    r.filter_before_save_check_box = None
    r.filter_pass_count_slider = None
    r.filter_spatial_sigma_slider = None
    r.filter_normal_angle_slider = None
    r._save_requested = True
    r._frozen_rgb9e5 = None
    r._release_frozen_buffer = False
    r._save_error = None
    r.reset_accumulator = False
    r.reset_texture_space = False
    r.texture_iteration = 3
    r.save_rgb9e5_button = None
    r.layout = SimpleNamespace(total_payload_count=123)
    return r


def test_filter_enabled_ping_pongs_then_packs_latest_buffer():
    r = _renderer()
    filter_calls, pack_calls = [], []
    r.surface_filter = SimpleNamespace(
        execute=lambda encoder, source, destination, **kw: filter_calls.append((source, destination)))
    r.rgb9e5_packer = SimpleNamespace(
        execute=lambda encoder, source, destination, count: pack_calls.append((source, destination, count)))
    working, scratch, packed = object(), object(), object()
    assert r._execute_rgb9e5_save(object(), working, scratch, packed)
    assert filter_calls == [(working, scratch), (scratch, working), (working, scratch), (scratch, working),
                            (working, scratch)]
    assert pack_calls == [(scratch, packed, 123)]
    assert r._frozen_rgb9e5 is packed
    assert r.reset_accumulator


def test_filter_disabled_packs_working_buffer_and_reset_discards_frozen():
    r = _renderer()
    r.filter_before_save_check_box = SimpleNamespace(value=False)
    r.surface_filter = None
    calls = []
    r.rgb9e5_packer = SimpleNamespace(
        execute=lambda encoder, source, destination, count: calls.append((source, destination, count)))
    working, packed = object(), object()
    assert r._execute_rgb9e5_save(object(), working, None, packed)
    assert calls == [(working, packed, 123)]
    r._discard_rgb9e5()
    assert r._frozen_rgb9e5 is None and r._release_frozen_buffer


def test_save_failure_resets_texture_space():
    r = _renderer()
    r.surface_filter = None

    def boom(*a):
        raise RuntimeError("pack failed")

    r.rgb9e5_packer = SimpleNamespace(execute=boom)
    assert not r._execute_rgb9e5_save(object(), object(), None, object())
    assert r._frozen_rgb9e5 is None and r._release_frozen_buffer
    assert r.reset_texture_space and r.texture_iteration == 0
    assert r._save_error == "pack failed"


def test_save_requires_an_iteration():
    r = TextureSpacePathTracingRenderer()
    r._request_rgb9e5_save()
    assert not r._save_requested
    assert "at least one texture iteration" in r._save_error
    r.texture_iteration = 1
    r._request_rgb9e5_save()
    assert r._save_requested


def test_camera_move_resets():
    r = TextureSpacePathTracingRenderer()
    r.texture_iteration = 5
    r.reset_accumulator = False
    r.reset_texture_space = False
    r.on_camera_move({"texture_space": False})
    assert r.reset_accumulator and not r.reset_texture_space and r.texture_iteration == 5
    r.on_camera_move({"texture_space": True})
    assert r.reset_texture_space and r.texture_iteration == 0 and r._release_frozen_buffer


def test_cache_state_follows_save_and_discard():
    r = _renderer()
    r.layout = SimpleNamespace(total_payload_count=123, total_surface_texels=100)
    st = r.cache_state()
    assert st["iterations"] == 3 and st["surface_slots"] == 100 and st["back_slots"] == 23
    assert not st["frozen"] and st["save_pending"] and st["cache_bytes"] == 123 * 16
    r.surface_filter = None
    r.filter_before_save_check_box = SimpleNamespace(value=False)
    r.rgb9e5_packer = SimpleNamespace(execute=lambda *a: None)
    assert r._execute_rgb9e5_save(object(), object(), None, object())
    st = r.cache_state()
    assert st["frozen"] and not st["save_pending"] and st["cache_bytes"] == 123 * 4
