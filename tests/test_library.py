"""The C-ABI library loads and exports every entry point its header declares.
No compute calls (CPU only)."""
import ctypes as ct
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = []
    for header in sorted((ROOT / 'include').glob('*.h')):
        text = re.sub(r'/\*.*?\*/', '', header.read_text(), flags=re.S)
        for m in re.finditer(r'^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(', text, flags=re.M):
            name = m.group(1)
            if name not in ('if', 'for', 'while', 'return', 'sizeof'):
                names.append(name)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ('tensor_apply', 'hex_engine_create', 'hex_apply', 'hex_cgs2', 'hex_last_error'):
        assert required in names


def test_library_exports_every_declared_symbol():
    from laqphysics_amd import native
    if not native.LIBRARY.exists():
        from laqphysics_amd import build
        build.build()
    lib = ct.CDLL(str(native.LIBRARY))
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert not missing, missing


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from laqphysics_amd import native
    monkeypatch.setattr(native, 'LIBRARY', tmp_path / 'absent.so')
    monkeypatch.setattr(native, '_lib', None)
    with pytest.raises(RuntimeError, match='no CPU fallback'):
        native.lib()
