"""The C-ABI library loads and exports every entry point its headers declare
(no compute call: these run without a GPU)."""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / 'include').glob('*.h'))


def declared_functions():
    names = []
    for h in HEADERS:
        text = re.sub(r'/\*.*?\*/', '', h.read_text(), flags=re.S)
        for m in re.finditer(r'^\s*(?:const\s+)?(?:int|void|double|char)\s*\*?\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(', text, re.M):
            names.append(m.group(1))
    return names


@pytest.fixture(scope='module')
def lib():
    from braids_hip import _lib, build
    if not _lib.LIB_PATH.exists():
        build.build()
    return _lib.load()


def test_headers_declare_entry_points():
    names = declared_functions()
    assert 'braids_create' in names and 'explicit_frequency' in names and 'braids_native_rates' in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    from braids_hip import _lib
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_invalid_arguments_fail_loudly_without_a_device(lib):
    import ctypes
    from braids_hip import _lib
    out = ctypes.c_void_p()
    rc = lib.braids_create(None, None, None, ctypes.byref(out))
    assert rc == -1 and b'null' in lib.braids_last_error()
    assert lib.braids_step(None, 0.1, 0.0, None) == -1
