"""C-ABI boundary checks that need no GPU: the library loads, exports every
entry point declared in include/peel355.h, and rejects bad arguments before
touching the device (positive return codes, include/peel355.h)."""

import ctypes
import re

import pytest

from conftest import ROOT


def declared():
    text = (ROOT / "include" / "peel355.h").read_text()
    return sorted(set(re.findall(r"\b((?:peel|pool_mlp)_[a-z_]+)\s*\(", text)))


def test_header_declares_the_reference_lifecycle():
    names = declared()
    for n in ("peel_create", "peel_warmup", "peel_enqueue", "peel_capture", "peel_launch",
              "peel_get_buffers", "peel_read_status", "peel_destroy", "peel_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from hip_peel import _lib

    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert set(_lib.EXPORTS) <= set(declared())
    assert lib.peel_version().decode().startswith("peel355")


def _desc(**kw):
    from hip_peel import _lib

    d = _lib.PeelDesc()
    d.num_vertices, d.num_faces = 0, 0
    d.camera, d.settings = 8, 8  # never dereferenced: validation fails first
    for f in ("frontier", "regions", "hits", "vertices_cap", "triangles", "rounds", "layer_capacity", "edge_vertices"):
        setattr(d, f, 16)
    for k, v in kw.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize(
    "change,code,message",
    [
        ({"regions": 0}, 4, "Capacities must be positive"),
        ({"edge_vertices": 129}, 5, "edge_vertices"),
        ({"layers": 17}, 6, "layers must be zero"),
        ({"route": 9}, 7, "unknown route"),
        ({"route": 1, "workers": 0, "pieces": 4}, 4, "Capacities must be positive"),
        ({"route": 1, "workers": 4, "pieces": 0}, 4, "Capacities must be positive"),
        ({"num_faces": -1}, 2, "out of range"),
        ({"camera": None}, 3, "null device array"),
    ],
)
def test_argument_errors(change, code, message):
    from hip_peel import _lib

    lib = _lib.load()
    handle = ctypes.c_void_p()
    r = lib.peel_create(ctypes.byref(_desc(**change)), ctypes.byref(handle))
    assert r == code
    assert message in lib.peel_last_error().decode()
    assert not handle.value
    with pytest.raises(ValueError, match=message):
        _lib.check(r)


def test_null_arguments():
    from hip_peel import _lib

    lib = _lib.load()
    assert lib.peel_create(None, None) == 1
    assert lib.peel_launch(None) == 1
    assert lib.peel_destroy(None) == 0


def _header_fields(struct):
    text = (ROOT / "include" / "peel355.h").read_text()
    body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (struct, struct), text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    return re.findall(r"([A-Za-z_][A-Za-z_0-9]*)\s*;", body)


@pytest.mark.parametrize("struct", ["PeelDesc", "PeelBuffers", "PeelOutputs"])
def test_ctypes_structs_follow_the_header(struct):
    """The ctypes binding lists the header's fields in the header's order."""
    from hip_peel import _lib

    assert [f[0] for f in getattr(_lib, struct)._fields_] == _header_fields(struct)
