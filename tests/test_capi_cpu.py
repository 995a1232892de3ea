"""C ABI checks that need no GPU: the library loads and exports every symbol
declared in include/hasty_amd.h; the Python mirror lists them all."""
import ctypes as C
import os
import re

import hastyrenderer_amd as ha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = []
    for fname in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fname.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fname)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"\b(hasty_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = _declared_symbols()
    assert "hasty_render_pt" in names and "hasty_trace_closest" in names
    assert set(names) == set(ha.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(ha.LIB_PATH)
    for name in _declared_symbols():
        assert hasattr(lib, name), name


def test_last_error_is_a_string():
    assert isinstance(ha.lib().hasty_last_error(), bytes)


def test_load_reports_missing_file():
    import pytest

    with pytest.raises(ha.HastyError, match="could not open"):
        ha.Scene("/nonexistent/renderJob.json")
