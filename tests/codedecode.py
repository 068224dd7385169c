"""A CPU reconstruction of a P-frame from its MB codes (fdvgpu_mbcode), built
from the oracle's block kernels: the decode half of pframe_decode
(fdv.h:6492-6737) restated over parsed codes, so that the host parser's
levels and the device decoder can both be checked against it.

Test infrastructure only."""
from __future__ import annotations

import ctypes as C

import numpy as np

import orclib as L


def _p(a):
    return C.c_void_p(a.ctypes.data)


def _padded(plane, w, h):
    buf, off, st = L.padded(np.ascontiguousarray(plane), w, h)
    return buf, off, st


def _add4(lev, pred, qp):
    lib = L.oracle()
    lv = np.ascontiguousarray(lev, np.int16)
    d = np.zeros(16, np.int32)
    r = np.zeros(16, np.int16)
    lib.orc_dequant4(_p(lv), _p(d), qp)
    lib.orc_idct4(_p(d), _p(r))
    return np.clip(pred.astype(int) + r.reshape(4, 4), 0, 255).astype(np.uint8)


def _add8(lev, pred, qp):
    lib = L.oracle()
    lv = np.ascontiguousarray(lev, np.int16)
    d = np.zeros(64, np.int32)
    r = np.zeros(64, np.int16)
    lib.orc_dequant8(_p(lv), _p(d), qp)
    lib.orc_idct8(_p(d), _p(r))
    return np.clip(pred.astype(int) + r.reshape(8, 8), 0, 255).astype(np.uint8)


def _residual(pred, lev, flags, n, qp):
    """decode_residual (fdv.h:5426) over an n x n prediction, regions raster."""
    out = pred.copy()
    for r in range(max(1, (n // 8) ** 2)):
        oy, ox = (r // max(1, n // 8)) * 8, (r % max(1, n // 8)) * 8
        fl = flags[r]
        if fl == 1:
            for b in range(4):
                y, x = oy + (b >> 1) * 4, ox + (b & 1) * 4
                out[y:y + 4, x:x + 4] = _add4(lev[64 * r + 16 * b:64 * r + 16 * b + 16], pred[y:y + 4, x:x + 4], qp)
        elif fl == 2:
            out[oy:oy + 8, ox:ox + 8] = _add8(lev[64 * r:64 * r + 64], pred[oy:oy + 8, ox:ox + 8], qp)
    return out


def reconstruct(codes, r0, r1, w, h, qp):
    """The decoded frame (tight I420, deblocked) of codes against references
    r0 (near) and r1 (far, or None)."""
    lib = L.oracle()
    cw, ch = w // 2, h // 2
    refs = []
    for r in (r0, r1 if r1 is not None else r0):
        y, u, v = L.split_i420(r, w, h)
        refs.append([_padded(y, w, h), _padded(u, cw, ch), _padded(v, cw, ch)])
    Y = np.zeros((h, w), np.uint8)
    U = np.zeros((ch, cw), np.uint8)
    V = np.zeros((ch, cw), np.uint8)
    mw = w // 16

    def mc(pl, ref, bx, by, n, mvx, mvy, chroma):
        buf, off, st = refs[ref][pl]
        d = np.zeros(n * n, np.uint8)
        fn = lib.orc_mc_chroma if chroma else lib.orc_mc_luma
        fn(C.c_void_p(buf.ctypes.data + off), st, bx, by, n, n, int(mvx), int(mvy), _p(d), n)
        return d.reshape(n, n)

    def nbs(P, bx, by, n):
        ht, hl = by > 0, bx > 0
        top = P[by - 1, bx:bx + n].copy() if ht else np.zeros(n, np.uint8)
        left = P[by:by + n, bx - 1].copy() if hl else np.zeros(n, np.uint8)
        tl = int(P[by - 1, bx - 1]) if ht and hl else 128
        return top, left, tl, int(ht), int(hl)

    def intra(P, bx, by, n, sub):
        top, left, tl, ht, hl = nbs(P, bx, by, n)
        pr = np.zeros(n * n, np.uint8)
        lib.orc_intra_nxn(int(sub), _p(top), _p(left), C.c_uint8(tl), n, ht, hl, _p(pr))
        return pr.reshape(n, n)

    for i, c in enumerate(codes):
        bx, by = (i % mw) * 16, (i // mw) * 16
        cbx, cby = bx // 2, by // 2
        mode = int(c["mode"])
        if mode == 2:
            Y[by:by + 16, bx:bx + 16] = _residual(intra(Y, bx, by, 16, c["sub"]), c["lev"], c["lflag"], 16, qp)
            for p, P in enumerate((U, V)):
                P[cby:cby + 8, cbx:cbx + 8] = _residual(intra(P, cbx, cby, 8, c["sub"]), c["clev"][64 * p:],
                                                        [c["cflag"][p]], 8, qp)
            continue
        if mode == 3:
            for q in range(4):
                qx, qy = (q & 1) * 8, (q >> 1) * 8
                pr = mc(0, 0, bx + qx, by + qy, 8, c["mv"][q][0], c["mv"][q][1], False)
                Y[by + qy:by + qy + 8, bx + qx:bx + qx + 8] = _residual(pr, c["lev"][64 * q:], [c["lflag"][q]], 8, qp)
            for p, P in enumerate((U, V)):
                for q in range(4):
                    qx, qy = (q & 1) * 4, (q >> 1) * 4
                    pr = mc(1 + p, 0, cbx + qx, cby + qy, 4, int(c["mv"][q][0] / 2), int(c["mv"][q][1] / 2), True)
                    P[cby + qy:cby + qy + 4, cbx + qx:cbx + qx + 4] = _add4(c["clev"][64 * p + 16 * q:], pr, qp)
            continue
        mvx, mvy, ref = int(c["mv"][0][0]), int(c["mv"][0][1]), int(c["ref"])
        pr = mc(0, ref, bx, by, 16, mvx, mvy, False)
        Y[by:by + 16, bx:bx + 16] = pr if mode == 0 else _residual(pr, c["lev"], c["lflag"], 16, qp)
        for p, P in enumerate((U, V)):
            cp = mc(1 + p, ref, cbx, cby, 8, int(mvx / 2), int(mvy / 2), True)
            if mode == 0 or not c["cflag"][p]:
                P[cby:cby + 8, cbx:cbx + 8] = cp
            else:
                P[cby:cby + 8, cbx:cbx + 8] = _residual(cp, c["clev"][64 * p:], [1], 8, qp)
    for P, pw, ph in ((Y, w, h), (U, cw, ch), (V, cw, ch)):
        lib.orc_deblock_plane(_p(P), pw, ph, pw, qp)
    return np.concatenate([Y.ravel(), U.ravel(), V.ravel()])
