"""The C ABI: libforkmatch.so loads and exports every function that
include/forkmatch.h declares (no device calls)."""
import ctypes
import re
from pathlib import Path

import forkmatch_amd as fm
from forkmatch_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "forkmatch.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_functions():
    names = declared_functions()
    assert "fm_match" in names and "fm_match_batch" in names and len(names) > 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_errors():
    assert fm._lib.lib().fm_version().startswith(b"forkmatch")
    try:
        fm.Query.load_g("/nonexistent/q.g")
    except fm.ForkMatchError as e:
        assert e.code == 232
    else:
        raise AssertionError("missing query file must raise")
