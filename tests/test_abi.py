"""The C ABI (include/*.h): the built library exports every declared entry
point, and the calls that need no device behave as documented without one."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fdv_amd import gpu as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("fdvgpu.h", "fdv_amd.h")]


def declared_functions(path):
    src = open(path).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"#.*", "", src)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", src)
    return sorted(set(n for n in names if n.startswith(("fdvgpu_", "fdvamd_"))))


def test_headers_declare_the_boundary():
    names = set()
    for h in HEADERS:
        names |= set(declared_functions(h))
    for must in ("fdvgpu_open", "fdvgpu_pframe", "fdvgpu_fetch", "fdvgpu_streams", "fdvgpu_recon",
                 "fdvgpu_decisions", "fdvgpu_set_given", "fdvamd_pframe_encode",
                 "fdvamd_pframe_batch", "fdvamd_pack_pframe"):
        assert must in names


@pytest.mark.parametrize("header", HEADERS, ids=os.path.basename)
def test_library_exports_every_declared_symbol(header):
    lib = G.lib()
    missing = [n for n in declared_functions(header) if not hasattr(lib, n)]
    assert not missing, f"declared in {os.path.basename(header)} but not exported: {missing}"


def test_mbinfo_layout_matches_header():
    # fdvgpu_mbinfo: 4 x u8 then int16 mv[4][2] -> 20 bytes, mv at offset 4
    assert G.MBINFO_DTYPE.itemsize == 20
    assert G.MBINFO_DTYPE.fields["mv"][1] == 4


def test_bad_arguments_return_error_codes():
    lib = G.lib()
    w = (C.c_int * 1)(100)          # not a multiple of 16
    h = (C.c_int * 1)(64)
    assert not lib.fdvgpu_open(1, w, h, 1)
    assert lib.fdvgpu_last_error()
    assert lib.fdvgpu_pframe(None, None, None, None, None, None, 0) != 0
    assert lib.fdvgpu_fetch(None) != 0
    lib.fdvamd_encoder_open.restype = C.c_void_p
    assert not lib.fdvamd_encoder_open(0, None, None, 0)


def test_pack_of_empty_frame_has_no_device_dependency():
    st = [b""] * 16
    st[0] = bytes(4)
    out = G.pack_pframe(20, st)
    assert len(out) > 0
    assert G.pack_pframe(20, st) == out
    assert isinstance(np.frombuffer(out, np.uint8), np.ndarray)


def _called_entry_points(path):
    src = open(path).read()
    return set(re.findall(r"\b(?:L|lib\(\))\.(fdv(?:gpu|amd)_[a-z0-9_]+)\(", src))


@pytest.mark.parametrize("path", ["fdv_amd/gpu.py", "fdv_amd/learned/train.py", "bench.py"])
def test_every_called_entry_point_has_a_signature(path):
    """A function called through ctypes without argtypes gets its handle
    narrowed to a C int: every entry point the Python side calls is declared."""
    lib = G.lib()
    called = _called_entry_points(os.path.join(ROOT, path))
    undeclared = [n for n in sorted(called) if getattr(lib, n).argtypes is None]
    assert not undeclared, f"{path} calls entry points without argtypes: {undeclared}"
    for n in G._signatures():
        assert hasattr(lib, n), f"signature table names {n}, which the library does not export"
