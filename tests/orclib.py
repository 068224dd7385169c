"""ctypes access to the parity oracle (oracle/liboracle.so) and, where it was
built, to the reference codec compiled from its own sources (oracle/_ref).

Test infrastructure only.  Both libraries are built by oracle/Makefile (also
run by __graft_entry__.build()).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfdvref.so")

NS = 16
u8p = C.POINTER(C.c_uint8)


def ptr(a: np.ndarray):
    return a.ctypes.data_as(u8p)


class MbInfo(C.Structure):
    _fields_ = [("mode", C.c_uint8), ("ref", C.c_uint8), ("merge_idx", C.c_uint8),
                ("sub", C.c_uint8), ("mv", (C.c_int16 * 2) * 4)]


TABCACHE_BYTES = 2 * 256 + 2 * 16 * 256 + 4 + 4 * 16 * 256 + 4


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_oracle = None
_ref = None


def oracle():
    global _oracle
    if _oracle is None:
        lib = _load(ORACLE_SO)
        if lib is None:
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        lib.orc_pframe_encode.restype = C.c_size_t
        lib.orc_pframe_pack.restype = C.c_size_t
        lib.orc_lambda.restype = C.c_double
        lib.orc_lambda.argtypes = [C.c_int, C.c_double]
        lib.orc_rdoq4.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double]
        lib.orc_code_residual.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                          C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_pframe_analyse.argtypes = [
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
            C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_void_p,
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
        lib.orc_pframe_encode.argtypes = [
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_int, C.c_void_p, C.c_double, C.c_int, C.c_void_p]
        lib.orc_pframe_pack.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_size_t]
        _oracle = lib
    return _oracle


def ref():
    """The reference build, or None when it was not built here."""
    global _ref
    if _ref is None:
        lib = _load(REF_SO)
        if lib is None:
            return None
        lib.fdvref_pframe_encode.restype = C.c_size_t
        lib.fdvref_pframe_encode.argtypes = [
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_int, C.c_int, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_int, C.c_void_p, C.c_double]
        lib.fdvref_pframe_decode.restype = C.c_int
        lib.fdvref_pframe_decode.argtypes = [
            C.c_void_p, C.c_size_t, C.c_int, C.c_int,
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.fdvref_tabcache_size.restype = C.c_size_t
        lib.fdvref_code_residual.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                             C.c_int, C.c_int, C.c_double, C.c_int, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.fdv_rdoq4x4.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double]
        lib.fdv_video_encode.restype = C.c_size_t
        lib.fdv_video_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_void_p, C.c_size_t]
        lib.fdv_video_decode.restype = C.c_int
        lib.fdv_video_decode.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p]
        _ref = lib
    return _ref


def split_i420(f: np.ndarray, w: int, h: int):
    ys = w * h
    cs = (w // 2) * (h // 2)
    return f[:ys], f[ys:ys + cs], f[ys + cs:ys + 2 * cs]


class TabCache:
    """Entropy state carried between P-frames (same layout in oracle and reference)."""

    def __init__(self):
        self.buf = np.zeros(TABCACHE_BYTES, dtype=np.uint8)

    def copy(self):
        t = TabCache()
        t.buf[:] = self.buf
        return t


def oracle_pframe(cur, r0, r1, w, h, navail, qp, tc=None, lscale=1.0, force_skip=0,
                  me_mode=0, want_info=False):
    """Oracle P-frame encode: returns (payload bytes, recon I420, mbinfo array)."""
    lib = oracle()
    cy, cu, cv = split_i420(cur, w, h)
    a = split_i420(r0, w, h)
    b = split_i420(r1, w, h) if r1 is not None else a
    cap = w * h * 3 + 65536
    out = np.zeros(cap, dtype=np.uint8)
    rec = np.zeros_like(cur)
    ry, ru, rv = split_i420(rec, w, h)
    nmb = (w // 16) * (h // 16)
    info = (MbInfo * nmb)()
    n = lib.orc_pframe_encode(cy.ctypes.data, cu.ctypes.data, cv.ctypes.data, w, h,
                              a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                              b[0].ctypes.data, b[1].ctypes.data, b[2].ctypes.data,
                              navail, qp, out.ctypes.data, cap,
                              ry.ctypes.data, ru.ctypes.data, rv.ctypes.data, force_skip,
                              tc.buf.ctypes.data if tc is not None else None, lscale, me_mode,
                              C.addressof(info))
    return out[:n].copy(), rec, info


def ref_pframe(cur, r0, r1, w, h, navail, qp, tc=None, lscale=1.0, force_skip=0):
    lib = ref()
    cy, cu, cv = split_i420(cur, w, h)
    a = split_i420(r0, w, h)
    b = split_i420(r1, w, h) if r1 is not None else a
    cap = w * h * 3 + 65536
    out = np.zeros(cap, dtype=np.uint8)
    rec = np.zeros_like(cur)
    ry, ru, rv = split_i420(rec, w, h)
    n = lib.fdvref_pframe_encode(cy.ctypes.data, cu.ctypes.data, cv.ctypes.data, w, h,
                                 a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                 b[0].ctypes.data, b[1].ctypes.data, b[2].ctypes.data,
                                 navail, qp, out.ctypes.data, cap,
                                 ry.ctypes.data, ru.ctypes.data, rv.ctypes.data, force_skip,
                                 tc.buf.ctypes.data if tc is not None else None, lscale)
    return out[:n].copy(), rec
