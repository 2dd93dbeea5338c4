"""The C-ABI library loads and exports every symbol include/nan_hip.h declares
(no compute calls: this runs without a GPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from nan_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def _declared_symbols() -> set[str]:
    text = (ROOT / "include" / "nan_hip.h").read_text()
    return set(re.findall(r"NAN_API\s+[\w\s\*]+?\b(nan_\w+)\s*\(", text))


def test_header_and_binding_agree():
    declared = _declared_symbols()
    assert declared == set(_native.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    if not _native.LIB_PATH.exists():
        pytest.skip("libnan_hip.so not built")
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    for name in sorted(_declared_symbols()):
        assert hasattr(lib, name), name
    assert _native.lib().nan_version() == _native.ABI_VERSION == 3


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, "LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(_native, "_LIB", None)
    with pytest.raises(_native.NativeLibraryError):
        _native.lib()


def test_oracle_is_not_imported_by_the_product():
    for path in (ROOT / "nan_amd").rglob("*.py"):
        text = path.read_text()
        assert "import oracle" not in text and "from oracle" not in text, path
