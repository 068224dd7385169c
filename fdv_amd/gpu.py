"""ctypes binding of libfdvgpu.so (include/fdvgpu.h): the MI355X encode-analysis
path.  There is no fallback: if the library is missing or the GPU call fails,
this raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfdvgpu.so")

ME_DIAMOND, ME_EXHAUSTIVE, ME_LEARNED = 0, 1, 2
NS = 16

MBINFO_DTYPE = np.dtype([("mode", "u1"), ("ref", "u1"), ("merge_idx", "u1"), ("sub", "u1"),
                         ("mv", "<i2", (4, 2))])

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("fdv_amd/libfdvgpu.so is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.fdvgpu_open.restype = vp
        L.fdvgpu_open.argtypes = [C.c_int, ip, ip, C.c_int]
        L.fdvgpu_close.argtypes = [vp]
        L.fdvgpu_last_error.restype = C.c_char_p
        L.fdvgpu_put_source.argtypes = [vp, C.c_int, C.c_int, vp]
        L.fdvgpu_put_ref.argtypes = [vp, C.c_int, C.c_int, vp]
        L.fdvgpu_set_given.argtypes = [vp, C.c_int, vp]
        L.fdvgpu_pframe.argtypes = [vp, ip, ip, dp, ip, ip, C.c_int]
        L.fdvgpu_fetch.argtypes = [vp]
        L.fdvgpu_streams.argtypes = [vp, C.c_int, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t)]
        L.fdvgpu_recon.argtypes = [vp, C.c_int, vp]
        L.fdvgpu_decisions.argtypes = [vp, C.c_int, vp]
        L.fdvgpu_timings.argtypes = [vp, C.POINTER(C.c_float)]
        _lib = L
    return _lib


def _ints(v):
    return (C.c_int * len(v))(*[int(x) for x in v])


class GpuError(RuntimeError):
    pass


class Analyser:
    """Batched P-frame analysis for a set of sequences (streams or bands)."""

    def __init__(self, sizes, nsrc: int = 1):
        self.L = lib()
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.nseq = len(self.sizes)
        ws = _ints([w for w, _ in self.sizes])
        hs = _ints([h for _, h in self.sizes])
        self.h = self.L.fdvgpu_open(self.nseq, ws, hs, int(nsrc))
        if not self.h:
            raise GpuError("fdvgpu_open: " + self.L.fdvgpu_last_error().decode())

    def _ck(self, rc, what):
        if rc != 0:
            raise GpuError(what + ": " + self.L.fdvgpu_last_error().decode())

    def close(self):
        if self.h:
            self.L.fdvgpu_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frame_bytes(self, seq):
        w, h = self.sizes[seq]
        return w * h + 2 * (w // 2) * (h // 2)

    def put_source(self, seq, slot, frame: np.ndarray):
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        assert f.size == self.frame_bytes(seq)
        self._ck(self.L.fdvgpu_put_source(self.h, seq, slot, f.ctypes.data), "put_source")

    def put_ref(self, seq, which, frame: np.ndarray):
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        assert f.size == self.frame_bytes(seq)
        self._ck(self.L.fdvgpu_put_ref(self.h, seq, which, f.ctypes.data), "put_ref")

    def set_given(self, seq, dec):
        if dec is None:
            self._ck(self.L.fdvgpu_set_given(self.h, seq, None), "set_given")
            return
        d = np.ascontiguousarray(dec, dtype=MBINFO_DTYPE)
        self._ck(self.L.fdvgpu_set_given(self.h, seq, d.ctypes.data), "set_given")

    def pframe(self, src_slot, qp, lscale=None, navail=None, force_skip=None, me_mode=ME_DIAMOND):
        n = self.nseq
        lscale = [1.0] * n if lscale is None else lscale
        navail = [2] * n if navail is None else navail
        force_skip = [0] * n if force_skip is None else force_skip
        ls = (C.c_double * n)(*[float(x) for x in lscale])
        self._ck(self.L.fdvgpu_pframe(self.h, _ints(src_slot), _ints(qp), ls, _ints(navail),
                                      _ints(force_skip), int(me_mode)), "pframe")

    def fetch(self):
        self._ck(self.L.fdvgpu_fetch(self.h), "fetch")

    def streams(self, seq):
        sv = (C.POINTER(C.c_uint8) * NS)()
        sn = (C.c_size_t * NS)()
        self._ck(self.L.fdvgpu_streams(self.h, seq, sv, sn), "streams")
        return [bytes(C.string_at(sv[k], sn[k])) if sn[k] else b"" for k in range(NS)]

    def recon(self, seq):
        out = np.zeros(self.frame_bytes(seq), dtype=np.uint8)
        self._ck(self.L.fdvgpu_recon(self.h, seq, out.ctypes.data), "recon")
        return out

    def decisions(self, seq):
        w, h = self.sizes[seq]
        out = np.zeros((w // 16) * (h // 16), dtype=MBINFO_DTYPE)
        self._ck(self.L.fdvgpu_decisions(self.h, seq, out.ctypes.data), "decisions")
        return out

    def timings(self):
        ms = (C.c_float * 4)()
        self._ck(self.L.fdvgpu_timings(self.h, ms), "timings")
        return list(ms)


TABCACHE_BYTES = 2 * 256 + 2 * 16 * 256 + 4 + 4 * 16 * 256 + 4


class TabCache:
    """The entropy state carried between P-frames (fdvamd_tabcache)."""

    def __init__(self):
        self.buf = np.zeros(TABCACHE_BYTES, dtype=np.uint8)


def pack_pframe(qp: int, streams, tc: TabCache | None = None) -> bytes:
    """Host entropy packing of one P-frame payload (fdvamd_pack_pframe)."""
    L = lib()
    L.fdvamd_pack_pframe.restype = C.c_size_t
    L.fdvamd_pack_pframe.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
    arrs = [np.frombuffer(s, dtype=np.uint8) if len(s) else np.zeros(1, dtype=np.uint8) for s in streams]
    sv = (C.c_void_p * NS)(*[a.ctypes.data for a in arrs])
    sn = (C.c_size_t * NS)(*[len(s) for s in streams])
    cap = sum(len(s) for s in streams) * 3 + 65536
    out = np.zeros(cap, dtype=np.uint8)
    n = L.fdvamd_pack_pframe(int(qp), C.addressof(sv), C.addressof(sn),
                             tc.buf.ctypes.data if tc is not None else None, out.ctypes.data, cap)
    if n == 0:
        raise GpuError("fdvamd_pack_pframe failed")
    return out[:n].tobytes()
