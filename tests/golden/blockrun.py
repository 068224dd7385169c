"""Runs the golden block-level cases through one of two CPU backends:

- "ref": the reference codec compiled from its own sources by oracle/Makefile
  (oracle/_ref/libfdvref.so; only where the reference tree is present), through
  its public block API (fdv.h:668-849) and the small shim in oracle/ref_shim.c;
- "oracle": the project's C restatement (oracle/liboracle.so).

Each runner returns one short digest per case, so a fixture written from the
reference and a check run against the oracle compare case by case.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import orclib as L  # noqa: E402
import vecgen as V  # noqa: E402

ME_RANGE = 16        # the reference's default search range in pels (fdv.h:5169)


def _p(a):
    return C.c_void_p(a.ctypes.data)


class Backend:
    def __init__(self, kind: str):
        self.kind = kind
        if kind == "ref":
            self.lib = L.ref()
            if self.lib is None:
                raise RuntimeError("reference build not present")
            n = {"fdct4": "fdv_fdct4x4", "idct4": "fdv_idct4x4", "quant4": "fdv_quant4x4",
                 "dequant4": "fdv_dequant4x4", "rdoq4": "fdv_rdoq4x4", "fdct8": "fdv_fdct8x8",
                 "idct8": "fdv_idct8x8", "quant8": "fdv_quant8x8", "dequant8": "fdv_dequant8x8",
                 "satd4": "fdv_satd4x4", "intra": "fdv_intra_nxn", "deblock": "fdv_deblock_plane",
                 "resid": "fdvref_code_residual"}
        else:
            self.lib = L.oracle()
            n = {"fdct4": "orc_fdct4", "idct4": "orc_idct4", "quant4": "orc_quant4",
                 "dequant4": "orc_dequant4", "rdoq4": "orc_rdoq4", "fdct8": "orc_fdct8",
                 "idct8": "orc_idct8", "quant8": "orc_quant8", "dequant8": "orc_dequant8",
                 "satd4": "orc_satd4", "intra": "orc_intra_nxn", "deblock": "orc_deblock_plane",
                 "resid": "orc_code_residual"}
        self.f = {k: getattr(self.lib, v) for k, v in n.items()}
        self.f["rdoq4"].argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double]
        self.f["resid"].argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                    C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]
        if kind == "ref":
            self.lib.fdvref_c8_tables.argtypes = [C.c_void_p, C.c_void_p]
            self.lib.fdvref_me_search.restype = C.c_int
        else:
            self.lib.orc_tables8.argtypes = [C.c_void_p, C.c_void_p]
            self.lib.orc_me_search.restype = C.c_int

    # -- tables -------------------------------------------------------------
    def tables8(self):
        c8 = np.zeros(64, np.int32)
        qs = np.zeros(52, np.int32)
        (self.lib.fdvref_c8_tables if self.kind == "ref" else self.lib.orc_tables8)(_p(c8), _p(qs))
        return c8, qs

    # -- 4x4 ----------------------------------------------------------------
    def transform4(self, res, qp):
        c = np.zeros(16, np.int32)
        lv = np.zeros(16, np.int16)
        d = np.zeros(16, np.int32)
        r = np.zeros(16, np.int16)
        self.f["fdct4"](_p(res), _p(c))
        self.f["quant4"](_p(c), _p(lv), qp)
        self.f["dequant4"](_p(lv), _p(d), qp)
        self.f["idct4"](_p(d), _p(r))
        # the inverse of the raw forward output as well (no quantiser in between)
        r2 = np.zeros(16, np.int16)
        self.f["idct4"](_p(c), _p(r2))
        return c, lv, d, r, r2

    def rdoq(self, res, qp, lam):
        c = np.zeros(16, np.int32)
        lv = np.zeros(16, np.int16)
        self.f["fdct4"](_p(res), _p(c))
        self.f["quant4"](_p(c), _p(lv), qp)
        self.f["rdoq4"](_p(lv), _p(res), qp, lam)
        return lv

    def transform8(self, x, qp):
        y = np.zeros(64, np.int32)
        lv = np.zeros(64, np.int16)
        d = np.zeros(64, np.int32)
        r = np.zeros(64, np.int16)
        self.f["fdct8"](_p(x), _p(y))
        self.f["quant8"](_p(y), _p(lv), qp)
        self.f["dequant8"](_p(lv), _p(d), qp)
        self.f["idct8"](_p(d), _p(r))
        return y, lv, d, r

    def satd(self, src, pred):
        return int(self.f["satd4"](_p(src), 4, _p(pred)))

    def intra(self, mode, n, ht, hl, top, left, tl):
        pred = np.zeros(n * n, np.uint8)
        self.f["intra"](mode, _p(top), _p(left), C.c_uint8(tl), n, ht, hl, _p(pred))
        return pred

    def deblock(self, img, w, h, qp):
        a = np.ascontiguousarray(img.copy())
        self.f["deblock"](_p(a), w, h, w, qp)
        return a

    def mc(self, plane, chroma, bx, by, bw, bh, mvx, mvy):
        h, w = plane.shape
        dst = np.zeros(bw * bh, np.uint8)
        if self.kind == "ref":
            self.lib.fdvref_mc(_p(np.ascontiguousarray(plane)), w, h, chroma, bx, by, bw, bh,
                               mvx, mvy, _p(dst))
        else:
            buf, off, st = L.padded(plane.ravel(), w, h)
            fn = self.lib.orc_mc_chroma if chroma else self.lib.orc_mc_luma
            fn(C.c_void_p(buf.ctypes.data + off), st, bx, by, bw, bh, mvx, mvy, _p(dst), bw)
        return dst

    def me(self, cur, ref, w, h, bx, by, bw, bh, pmx, pmy):
        mx, my = C.c_int(), C.c_int()
        cur = np.ascontiguousarray(cur)
        if self.kind == "ref":
            s = self.lib.fdvref_me_search(_p(cur), w, h, _p(np.ascontiguousarray(ref)), bx, by, bw, bh,
                                          ME_RANGE, pmx, pmy, C.byref(mx), C.byref(my))
        else:
            buf, off, st = L.padded(ref.ravel(), w, h)
            s = self.lib.orc_me_search(_p(cur), w, C.c_void_p(buf.ctypes.data + off), st, bx, by, bw, bh,
                                       ME_RANGE, pmx, pmy, 0, C.byref(mx), C.byref(my))
        return int(mx.value), int(my.value), int(s)

    def residual(self, cur, pred, n, qp, lam, chroma):
        cap = 4096
        sb = np.zeros(7 * cap, np.uint8)
        bufs = (C.c_void_p * 7)(*[sb.ctypes.data + k * cap for k in range(7)])
        lens = (C.c_size_t * 7)()
        rec = np.zeros(n * n, np.uint8)
        ssd = C.c_double()
        bits = C.c_int()
        self.f["resid"](_p(np.ascontiguousarray(cur)), n, 0, 0, _p(np.ascontiguousarray(pred)), n, qp,
                        lam, chroma, _p(rec), C.addressof(bufs), C.addressof(lens),
                        C.byref(ssd), C.byref(bits))
        streams = b"".join(bytes([k]) + len(sb[k * cap:k * cap + lens[k]]).to_bytes(4, "little") +
                           sb[k * cap:k * cap + lens[k]].tobytes() for k in range(7))
        return streams, rec, float(ssd.value), int(bits.value)


def run_blocks(kind: str) -> dict:
    """All block-level cases through one backend: {group: [per-case digest]}."""
    B = Backend(kind)
    out = {}
    c8, qs = B.tables8()
    out["tables8"] = {"c8": c8.tolist(), "qstep8": qs.tolist()}
    out["transform4"] = [V.short(*B.transform4(r, q)) for r, q in V.transform4_cases()]
    out["rdoq4"] = [V.short(B.rdoq(r, q, lam)) for r, q, lam in V.rdoq_cases()]
    out["transform8"] = [V.short(*B.transform8(x, q)) for x, q in V.transform8_cases()]
    out["satd4"] = [B.satd(a, b) for a, b in V.satd_cases()]
    out["intra"] = [V.short(B.intra(*c)) for c in V.intra_cases()]
    plane, cases = V.mc_cases()
    out["mc"] = [V.short(B.mc(plane, *c)) for c in cases]
    out["me"] = [list(B.me(*c)) for c in V.me_cases()]
    out["deblock"] = [V.short(B.deblock(*c)) for c in V.deblock_cases()]
    res = []
    for cur, pred, n, qp, lam, chroma in V.residual_cases():
        st, rec, ssd, bits = B.residual(cur, pred, n, qp, lam, chroma)
        res.append([V.short(st, rec), repr(ssd), bits])
    out["residual"] = res
    return out


def run_chains(kind: str) -> list:
    """Whole P-frame chains (payload bytes and reconstruction per frame) through
    the reference encoder or the oracle's analyse + pack."""
    out = []
    for name, spec, nfr, qp, lscale, fs_at in V.pframe_chain_specs():
        sc = V.build_scene(spec)
        w, h = sc.w, sc.h
        frames = [sc.frame(i) for i in range(nfr)]
        tc = L.TabCache()
        near, far, nref = frames[0], None, 1
        recs = []
        for f in range(1, nfr):
            navail = 2 if nref >= 2 else 1
            fs = 1 if fs_at == f else 0
            if kind == "ref":
                pay, rec = L.ref_pframe(frames[f], near, far, w, h, navail, qp, tc, lscale, fs)
                pay = pay.tobytes()
            else:
                st, rec, _ = L.oracle_analyse(frames[f], near, far, w, h, navail, qp, lscale, fs)
                pay = L.oracle_pack(qp, st, tc)
            recs.append({"frame": f, "input": V.short(frames[f]), "payload_len": len(pay),
                         "payload": V.digest(pay), "recon": V.digest(rec)})
            far, near, nref = near, rec, min(nref + 1, 2)
        out.append({"name": name, "w": w, "h": h, "qp": qp, "lscale": lscale,
                    "ref_input": V.short(frames[0]), "frames": recs})
    return out
