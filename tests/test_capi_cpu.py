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


def _tiny_job(tmp_path, settings, obj_faces="f 1 2 3\n", extra_obj=""):
    import json

    (tmp_path / "m.mtl").write_text("newmtl w\nKd 0.5 0.5 0.5\nNs 0\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\no t\nusemtl w\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\n" +
                                    extra_obj + obj_faces)
    json.dump({"camera": {"position": [0, 0, 3]}, "world": {"background_color": [1, 1, 1]}, "mesh": "m.obj"},
              open(tmp_path / "scene.json", "w"))
    json.dump({"settings": settings, "scene": "scene.json"}, open(tmp_path / "job.json", "w"))
    return str(tmp_path / "job.json")


def test_load_rejects_max_depth_above_255(tmp_path):
    """The wavefront path state keeps the depth in 8 bits: a job asking for
    more is refused at load (before any device work)."""
    import pytest

    with pytest.raises(ha.HastyError, match="max_depth above 255"):
        ha.Scene(_tiny_job(tmp_path, {"width": 4, "height": 4, "max_depth": 256, "use_mis": False}))


def test_loader_rejects_out_of_range_normal_and_texcoord_indices(tmp_path):
    """'f 1//9' with one vn line, or 'f 1/4' with one vt line, is a malformed
    asset (TinyObjReader error), for the library and the oracle alike."""
    import pytest

    from oracle import pyoracle as po

    st = {"width": 4, "height": 4, "use_mis": False}
    for faces, what in (("f 1//1 2//9 3//1\n", "normal"), ("f 1/1 2/4 3/1\n", "texcoord")):
        job = _tiny_job(tmp_path, st, obj_faces=faces, extra_obj="vt 0 0\n")
        with pytest.raises(ha.HastyError, match="invalid %s index" % what):
            ha.Scene(job)
        with pytest.raises(RuntimeError, match="invalid %s index" % what):
            po.OracleJob(job)
    # in range: loads in the oracle
    po.OracleJob(_tiny_job(tmp_path, st, obj_faces="f 1/1/1 2/1/1 3/1/1\n", extra_obj="vt 0 0\n"))


def test_bindings_reject_wrong_films():
    """Shape, dtype and contiguity are checked with ValueError (not assert)
    before a pointer reaches the C side."""
    import numpy as np
    import pytest

    with pytest.raises(ValueError):
        ha._fptr(np.zeros(4, np.float64))
    with pytest.raises(ValueError):
        ha._fptr(np.zeros((4, 4), np.float32)[:, ::2])
    with pytest.raises(ValueError):
        ha._check_films((np.zeros((2, 3, 3), np.float32),) * 3, 3, 2)
