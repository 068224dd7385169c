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
        lib.orc_image_encode.restype = C.c_size_t
        lib.orc_image_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                         C.c_size_t, C.c_void_p]
        lib.orc_iframe_encode.restype = C.c_size_t
        lib.orc_iframe_encode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                          C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_intra4.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_uint8, C.c_int, C.c_int, C.c_void_p]
        lib.orc_itree_alloc.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.orc_itree_free.argtypes = [C.c_void_p]
        lib.orc_image_analyse.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p]
        lib.orc_image_pack.restype = C.c_size_t
        lib.orc_image_pack.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_size_t]
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
        lib.fdv_vtile_encode.restype = C.c_size_t
        lib.fdv_vtile_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_void_p, C.c_size_t]
        lib.fdv_vtile_decode.restype = C.c_int
        lib.fdv_vtile_decode.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_int]
        lib.fdv_image_encode.restype = C.c_size_t
        lib.fdv_image_encode.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                         C.c_size_t, C.c_void_p]
        lib.fdvref_iframe_encode.restype = C.c_size_t
        lib.fdvref_iframe_encode.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                             C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p, C.c_void_p]
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


def padded(plane: np.ndarray, w: int, h: int):
    """A bordered copy of a tight plane in the reference layout (fdv.h:1548):
    stride align32(w + 64), 32 replicated samples each side.  Returns (buffer,
    byte offset of pixel (0,0), stride)."""
    stride = (w + 64 + 31) & ~31
    buf = np.zeros((h + 64) * stride, dtype=np.uint8)
    p = np.pad(plane.reshape(h, w), 32, mode="edge")
    buf.reshape(h + 64, stride)[:, :w + 64] = p
    return buf, 32 * stride + 32, stride


def oracle_analyse(cur, r0, r1, w, h, navail, qp, lscale=1.0, force_skip=0, me_mode=0, given=None):
    """Oracle analysis of one P-frame: (16 streams as bytes, deblocked recon I420,
    per-MB MbInfo array).  `given` is an MbInfo array or None."""
    lib = oracle()
    cy, cu, cv = split_i420(cur, w, h)
    keep = []
    ptrs = [[0, 0], [0, 0], [0, 0]]
    refs = [r0, r1 if r1 is not None else r0]
    ys = cs = 0
    for r in range(2):
        pl = split_i420(refs[r], w, h)
        for k, (p, pw, ph) in enumerate(((pl[0], w, h), (pl[1], w // 2, h // 2), (pl[2], w // 2, h // 2))):
            buf, off, st = padded(p, pw, ph)
            keep.append(buf)
            ptrs[k][r] = buf.ctypes.data + off
            if k == 0:
                ys = st
            else:
                cs = st
    RY = (C.c_void_p * 2)(*ptrs[0])
    RU = (C.c_void_p * 2)(*ptrs[1])
    RV = (C.c_void_p * 2)(*ptrs[2])
    nmb = (w // 16) * (h // 16)
    scap = nmb * 1600 + 1024
    sb = np.zeros(NS * scap, dtype=np.uint8)
    bufs = (C.c_void_p * NS)(*[sb.ctypes.data + k * scap for k in range(NS)])
    sn = (C.c_size_t * NS)()
    rec = np.zeros_like(cur)
    ry, ru, rv = split_i420(rec, w, h)
    info = (MbInfo * nmb)()
    lam = lib.orc_lambda(qp, lscale)
    rc = lib.orc_pframe_analyse(cy.ctypes.data, cu.ctypes.data, cv.ctypes.data, w, h,
                                C.addressof(RY), C.addressof(RU), C.addressof(RV), ys, cs,
                                navail, qp, lam, force_skip, me_mode,
                                C.addressof(given) if given is not None else None,
                                ry.ctypes.data, ru.ctypes.data, rv.ctypes.data, C.addressof(info),
                                C.addressof(bufs), scap, C.addressof(sn))
    assert rc == 0, "oracle stream overflow"
    streams = [bytes(sb[k * scap:k * scap + sn[k]]) for k in range(NS)]
    return streams, rec, info


def oracle_pack(qp, streams, tc=None):
    lib = oracle()
    arrs = [np.frombuffer(s, dtype=np.uint8) if s else np.zeros(1, dtype=np.uint8) for s in streams]
    sv = (C.c_void_p * NS)(*[a.ctypes.data for a in arrs])
    sn = (C.c_size_t * NS)(*[len(s) for s in streams])
    total = sum(len(s) for s in streams)
    cap = total * 3 + 65536
    out = np.zeros(cap, dtype=np.uint8)
    n = lib.orc_pframe_pack(qp, C.addressof(sv), C.addressof(sn),
                            tc.buf.ctypes.data if tc is not None else None, out.ctypes.data, cap)
    return out[:n].tobytes()


class Metab(C.Structure):
    """orc_metab / fdvg_metab: per-MB motion table."""
    _fields_ = [("mv16", (C.c_int16 * 2) * 2), ("sad16", C.c_int32 * 2), ("mv8", (C.c_int16 * 2) * 4),
                ("sad8", C.c_int32 * 4), ("flags", C.c_int32)]


def _padded_refs(r0, r1, w, h):
    keep, ptrs = [], [[0, 0], [0, 0], [0, 0]]
    refs = [r0, r1 if r1 is not None else r0]
    ys = cs = 0
    for r in range(2):
        pl = split_i420(refs[r], w, h)
        for k, (p, pw, ph) in enumerate(((pl[0], w, h), (pl[1], w // 2, h // 2), (pl[2], w // 2, h // 2))):
            buf, off, st = padded(p, pw, ph)
            keep.append(buf)
            ptrs[k][r] = buf.ctypes.data + off
            if k == 0:
                ys = st
            else:
                cs = st
    return keep, ptrs, ys, cs


def oracle_metab_exhaustive(cur, r0, r1, w, h, navail):
    lib = oracle()
    keep, ptrs, ys, cs = _padded_refs(r0, r1, w, h)
    RY = (C.c_void_p * 2)(*ptrs[0])
    nmb = (w // 16) * (h // 16)
    mt = (Metab * nmb)()
    cy = split_i420(cur, w, h)[0]
    lib.orc_metab_exhaustive.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    lib.orc_metab_exhaustive(C.c_void_p(cy.ctypes.data), w, h, C.addressof(RY), ys, navail, C.addressof(mt))
    return mt


def oracle_parallel(cur, r0, r1, w, h, navail, qp, lscale=1.0, metab=None):
    """The block-parallel modes' result: pframe_encode's MB loop on a motion
    table (orc_pframe_analyse_table), the exhaustive table when none is given.
    Returns (16 streams as bytes, deblocked recon I420, per-MB MbInfo array)."""
    lib = oracle()
    keep, ptrs, ys, cs = _padded_refs(r0, r1, w, h)
    RY = (C.c_void_p * 2)(*ptrs[0])
    RU = (C.c_void_p * 2)(*ptrs[1])
    RV = (C.c_void_p * 2)(*ptrs[2])
    nmb = (w // 16) * (h // 16)
    mt = metab if metab is not None else oracle_metab_exhaustive(cur, r0, r1, w, h, navail)
    cy, cu, cv = split_i420(cur, w, h)
    scap = nmb * 1600 + 1024
    sb = np.zeros(NS * scap, dtype=np.uint8)
    bufs = (C.c_void_p * NS)(*[sb.ctypes.data + k * scap for k in range(NS)])
    sn = (C.c_size_t * NS)()
    rec = np.zeros_like(cur)
    ry, ru, rv = split_i420(rec, w, h)
    info = (MbInfo * nmb)()
    lam = lib.orc_lambda(qp, lscale)
    lib.orc_pframe_analyse_table.argtypes = [C.c_void_p] * 3 + [C.c_int] * 2 + [C.c_void_p] * 3 + \
        [C.c_int] * 4 + [C.c_double, C.c_void_p] + [C.c_void_p] * 4 + [C.c_void_p, C.c_size_t, C.c_void_p]
    rc = lib.orc_pframe_analyse_table(cy.ctypes.data, cu.ctypes.data, cv.ctypes.data, w, h,
                                      C.addressof(RY), C.addressof(RU), C.addressof(RV), ys, cs, navail, qp, lam,
                                      C.addressof(mt), ry.ctypes.data, ru.ctypes.data, rv.ctypes.data,
                                      C.addressof(info), C.addressof(bufs), scap, C.addressof(sn))
    assert rc == 0, "oracle stream overflow"
    streams = [bytes(sb[k * scap:k * scap + sn[k]]) for k in range(NS)]
    return streams, rec, info


def set_no_deblock(on: bool):
    oracle().orc_set_no_deblock(1 if on else 0)


def oracle_metab_candidates(cur, r0, r1, w, h, navail, nnout, thr=0.7, sthr=2.0):
    """The learned mode's motion table from the proposer outputs `nnout` (a numpy
    array with the fdvgpu_proposal layout)."""
    lib = oracle()
    keep, ptrs, ys, cs = _padded_refs(r0, r1, w, h)
    RY = (C.c_void_p * 2)(*ptrs[0])
    nmb = (w // 16) * (h // 16)
    mt = (Metab * nmb)()
    cy = split_i420(cur, w, h)[0]
    nn = np.ascontiguousarray(nnout)
    assert nn.dtype.itemsize == 48 and nn.size == nmb
    lib.orc_metab_candidates.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p,
                                         C.c_float, C.c_float, C.c_void_p]
    lib.orc_metab_candidates(C.c_void_p(cy.ctypes.data), w, h, C.addressof(RY), ys, navail,
                             C.c_void_p(nn.ctypes.data), float(thr), float(sthr), C.addressof(mt))
    return mt


def metab_to_np(mt):
    import numpy as _np
    dt = _np.dtype([("mv16", "<i2", (2, 2)), ("sad16", "<i4", 2), ("mv8", "<i2", (4, 2)),
                    ("sad8", "<i4", 4), ("flags", "<i4")])
    return _np.frombuffer(bytes(mt), dtype=dt).copy()


def ref_pframe_decode(pay, r0, r1, w, h, navail, tc=None):
    """The reference decoder (pframe_decode, fdv.h:6492) on one P-frame payload:
    the decoded I420 frame, or None if it rejects the payload."""
    lib = ref()
    pay = np.ascontiguousarray(np.frombuffer(bytes(pay), dtype=np.uint8))
    a = split_i420(r0, w, h)
    b = split_i420(r1 if r1 is not None else r0, w, h)
    out = np.zeros(w * h + 2 * (w // 2) * (h // 2), dtype=np.uint8)
    oy, ou, ov = split_i420(out, w, h)
    rc = lib.fdvref_pframe_decode(pay.ctypes.data, len(pay), w, h,
                                  a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                  b[0].ctypes.data, b[1].ctypes.data, b[2].ctypes.data,
                                  navail, oy.ctypes.data, ou.ctypes.data, ov.ctypes.data,
                                  tc.buf.ctypes.data if tc is not None else None)
    return out if rc == 0 else None


def ref_video_encode(frames, w, h, qp, keyint, band_mbrows=0, nthreads=8):
    """The reference's fdv_video_encode (or, with band_mbrows, fdv_vtile_encode)."""
    lib = ref()
    fr = [np.ascontiguousarray(f, dtype=np.uint8) for f in frames]
    fp = (C.c_void_p * len(fr))(*[f.ctypes.data for f in fr])
    cap = len(fr) * (w * h * 3 + 8192) + 65536
    out = np.zeros(cap, np.uint8)
    if band_mbrows:
        n = lib.fdv_vtile_encode(C.addressof(fp), len(fr), w, h, qp, keyint, band_mbrows, nthreads,
                                 out.ctypes.data, cap)
    else:
        n = lib.fdv_video_encode(C.addressof(fp), len(fr), w, h, qp, keyint, out.ctypes.data, cap)
    return out[:n].tobytes()


def ref_video_decode(stream, tiled=False, nthreads=8):
    """The reference's fdv_video_decode / fdv_vtile_decode: (frames array
    [n, w*h*3/2], w, h), or None if it rejects the stream."""
    lib = ref()
    data = np.frombuffer(stream, dtype=np.uint8).copy()
    b = bytes(stream[:8])
    if tiled:
        w, h, n = b[0] | b[1] << 8, b[2] | b[3] << 8, b[4] | b[5] << 8
    else:
        n, w, h = b[0] | b[1] << 8, b[2] | b[3] << 8, b[4] | b[5] << 8
    out = np.zeros(n * (w * h * 3 // 2), np.uint8)
    nf, ww, hh = C.c_int(), C.c_int(), C.c_int()
    if tiled:
        rc = lib.fdv_vtile_decode(data.ctypes.data, len(data), out.ctypes.data, C.byref(nf), C.byref(ww),
                                  C.byref(hh), nthreads)
    else:
        rc = lib.fdv_video_decode(data.ctypes.data, len(data), out.ctypes.data, C.byref(nf), C.byref(ww),
                                  C.byref(hh))
    if rc != 0:
        return None
    return out.reshape(n, -1), w, h


def ref_iframe_coder():
    """Function pointers (ints) of the reference's key-frame coder and decoder
    (iframe_encode / iframe_decode through oracle/ref_shim.c)."""
    lib = ref()
    return (C.cast(lib.fdvref_iframe_encode, C.c_void_p).value, C.cast(lib.fdvref_iframe_decode, C.c_void_p).value)


class ITree(C.Structure):
    """orc_itree: a plane's decided intra quadtree in position-indexed planes."""
    _fields_ = [("rec", C.c_void_p), ("lev", C.c_void_p), ("rflag", C.c_void_p), ("cmode", C.c_void_p),
                ("split", C.c_void_p * 4), ("nmode", C.c_void_p * 4)]


def image_encode(lib, fn, src: np.ndarray, qp: int):
    """fdv_image_encode (or the oracle's) on a tight plane: (bytes, recon)."""
    h, w = src.shape
    src = np.ascontiguousarray(src, dtype=np.uint8)
    cap = w * h * 6 + 8192
    out = np.zeros(cap, np.uint8)
    rec = np.zeros((h, w), np.uint8)
    n = getattr(lib, fn)(src.ctypes.data, w, h, w, qp, out.ctypes.data, cap, rec.ctypes.data)
    return out[:n].copy(), rec


def image_tree(src: np.ndarray, qp: int):
    """The oracle's decided tree of a plane, as numpy arrays."""
    lib = oracle()
    h, w = src.shape
    src = np.ascontiguousarray(src, dtype=np.uint8)
    t = ITree()
    if lib.orc_itree_alloc(C.byref(t), w, h) != 0:
        raise MemoryError
    try:
        lib.orc_image_analyse(src.ctypes.data, w, h, w, qp, C.byref(t))

        def arr(p, shape, dt):
            n = int(np.prod(shape))
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,)).reshape(shape).copy()
        out = {"rec": arr(t.rec, (h, w), np.uint8), "lev": arr(t.lev, (h, w), np.int16),
               "rflag": arr(t.rflag, ((h + 7) // 8, (w + 7) // 8), np.uint8),
               "cmode": arr(t.cmode, (h // 4, w // 4), np.uint8)}
        for d in range(4):
            n = 64 >> d
            shp = ((h + n - 1) // n, (w + n - 1) // n)
            out["split%d" % d] = arr(t.split[d], shp, np.uint8)
            out["nmode%d" % d] = arr(t.nmode[d], shp, np.uint8)
        return out
    finally:
        lib.orc_itree_free(C.byref(t))
