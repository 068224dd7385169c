"""Deterministic inputs for the golden block-level vectors.

Both the fixture writer (make_golden.py, which runs the reference build) and the
CPU tests (which run the oracle) draw their inputs from here, so a fixture only
has to hold the expected outputs.  The generator is a self-contained splitmix64,
so the inputs do not depend on numpy's random streams.
"""
from __future__ import annotations

import hashlib

import numpy as np

M64 = (1 << 64) - 1


class Rng:
    def __init__(self, seed: int):
        self.s = seed & M64

    def u64(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def int(self, lo: int, hi: int) -> int:
        """Uniform in [lo, hi]."""
        return lo + self.u64() % (hi - lo + 1)

    def ints(self, n: int, lo: int, hi: int) -> np.ndarray:
        return np.array([self.int(lo, hi) for _ in range(n)], dtype=np.int64)

    def bytes_smooth(self, h: int, w: int, amp: int = 40) -> np.ndarray:
        """A smooth-ish u8 image: a bilinear ramp plus bounded noise."""
        a, b, c = self.int(-6, 6), self.int(-6, 6), self.int(40, 215)
        y, x = np.mgrid[0:h, 0:w]
        base = c + (a * x) // 4 + (b * y) // 4
        noise = self.ints(h * w, -amp, amp).reshape(h, w)
        return np.clip(base + noise, 0, 255).astype(np.uint8)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        if isinstance(a, (bytes, bytearray)):
            h.update(a)
        else:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def short(*arrays) -> str:
    return digest(*arrays)[:16]


# ---- case lists -----------------------------------------------------------

def transform4_cases(n=96):
    r = Rng(0x4A11)
    out = []
    for i in range(n):
        amp = [4, 32, 255, 510][i % 4]
        res = r.ints(16, -amp, amp).astype(np.int16)
        out.append((res, r.int(0, 51)))
    return out


def rdoq_cases(n=400):
    """(levels after quant4 of a residual, that residual, qp, lambda)."""
    r = Rng(0x7D09)
    out = []
    for i in range(n):
        amp = [3, 12, 40, 120, 255][i % 5]
        res = r.ints(16, -amp, amp)
        if i % 3 == 0:                     # sparse: energy in the low frequencies
            res = res * (np.arange(16) < r.int(1, 16))
        qp = r.int(0, 51)
        lam = [0.0, 0.425, 3.7, 27.0, 210.0][r.int(0, 4)] * (1.0 + r.int(0, 99) / 100.0)
        out.append((res.astype(np.int16), qp, float(lam)))
    return out


def transform8_cases(n=64):
    r = Rng(0x8E8E)
    out = []
    for i in range(n):
        amp = [4, 64, 255, 510][i % 4]
        out.append((r.ints(64, -amp, amp).astype(np.int16), r.int(0, 51)))
    return out


def satd_cases(n=64):
    r = Rng(0x5A7D)
    return [(r.ints(16, 0, 255).astype(np.uint8), r.ints(16, 0, 255).astype(np.uint8))
            for _ in range(n)]


def intra_cases():
    r = Rng(0x1A7A)
    out = []
    for n in (4, 8, 16):
        for mode in range(4):
            for ht in (0, 1):
                for hl in (0, 1):
                    for _ in range(3):
                        top = r.ints(16, 0, 255).astype(np.uint8)
                        left = r.ints(16, 0, 255).astype(np.uint8)
                        out.append((mode, n, ht, hl, top, left, r.int(0, 255)))
    return out


def mc_cases():
    """(plane h x w, chroma, bx, by, bw, bh, mvx, mvy) covering every phase,
    block sizes 4/8/16 and vectors that reach into the replicated border."""
    r = Rng(0x3C3C)
    w, h = 48, 32
    plane = r.bytes_smooth(h, w, amp=60)
    out = []
    for chroma in (0, 1):
        for fy in range(4):
            for fx in range(4):
                for bs in ((16, 16), (8, 8), (4, 4), (16, 8)):
                    bw, bh = bs
                    bx = r.int(0, (w - bw) // 4) * 4
                    by = r.int(0, (h - bh) // 4) * 4
                    ix, iy = r.int(-20, 20), r.int(-20, 20)
                    out.append((chroma, bx, by, bw, bh, ix * 4 + fx, iy * 4 + fy))
    return plane, out


def me_cases():
    """Small frames whose content is a shifted copy of the reference plus noise.
    Returns a list of (cur, ref, w, h, bx, by, bw, bh, pmx, pmy)."""
    r = Rng(0x3E3E)
    out = []
    for i in range(48):
        w, h = 64, 48
        big = r.bytes_smooth(h + 32, w + 32, amp=[3, 20, 50][i % 3])
        dx, dy = r.int(-12, 12), r.int(-10, 10)
        ref = big[16:16 + h, 16:16 + w].copy()
        cur = big[16 + dy:16 + dy + h, 16 + dx:16 + dx + w].copy()
        if i % 4 == 3:
            cur = np.clip(cur.astype(np.int64) + r.ints(h * w, -4, 4).reshape(h, w), 0, 255).astype(np.uint8)
        bw = bh = 16 if i % 2 == 0 else 8
        bx = r.int(0, (w - bw) // 8) * 8
        by = r.int(0, (h - bh) // 8) * 8
        pmx, pmy = r.int(-70, 70), r.int(-70, 70)
        out.append((cur, ref, w, h, bx, by, bw, bh, pmx, pmy))
    return out


def deblock_cases():
    r = Rng(0xDEB1)
    out = []
    for w in (16, 20, 64, 132, 320):
        h = 24 if w != 320 else 16
        for qp in (0, 7, 18, 24, 31, 40, 51):
            img = r.bytes_smooth(h, w, amp=r.int(2, 30))
            out.append((img, w, h, qp))
    return out


def residual_cases():
    """(cur n x n, pred n x n, n, qp, lambda, chroma) for code_residual."""
    r = Rng(0xC0DE)
    out = []
    for n in (4, 8, 16):
        for chroma in (0, 1):
            if n == 16 and chroma:
                continue
            for k in range(12):
                qp = [0, 8, 16, 24, 28, 32, 40, 51][k % 8]
                lam = 0.425 * 2.0 ** ((qp - 12) / 3.0) * [0.6, 1.0, 1.7][k % 3]
                cur = r.bytes_smooth(n, n, amp=[2, 10, 40, 90][k % 4])
                pred = np.clip(cur.astype(np.int64) + r.ints(n * n, -[1, 6, 30, 80][(k + 1) % 4],
                                                             [1, 6, 30, 80][(k + 1) % 4]).reshape(n, n),
                               0, 255).astype(np.uint8)
                out.append((cur, pred, n, qp, lam, chroma))
    return out


def pframe_chain_specs():
    """P-frame chains for whole-frame fixtures: (name, scene builder args, nframes,
    qp, lscale, force_skip_at).  Frame 0 is the reference; each later frame is
    coded against the previous reconstruction(s), the entropy state carried."""
    return [
        ("objects64x48_q28", ("objects", 64, 48, 3, {}), 5, 28, 1.0, None),
        ("pan96x64_q20", ("pan", 96, 64, 5, dict(nobj=3, smin=8, smax=32)), 5, 20, 1.0, None),
        ("pan96x64_q40_l0.8", ("pan", 96, 64, 6, dict(nobj=3, smin=8, smax=32)), 4, 40, 0.8, None),
        ("objects128x64_q12", ("objects", 128, 64, 8, dict(nobj=5, smin=8, smax=40, vmax=6, grain=2.0)), 4, 12, 1.0, None),
        ("objects80x48_q51", ("objects", 80, 48, 9, dict(nobj=2, smin=8, smax=24)), 4, 51, 1.0, None),
        ("objects80x48_q0", ("objects", 80, 48, 10, dict(nobj=2, smin=8, smax=24)), 3, 0, 1.0, None),
        ("pan64x64_q24_skip", ("pan", 64, 64, 11, dict(nobj=2, smin=8, smax=24)), 4, 24, 1.0, 2),
        ("config1_320x192_q28", ("objects", 320, 180, 1, dict(nobj=4, smin=24, smax=64, vmax=8)), 4, 28, 1.0, None),
    ]


def build_scene(spec):
    from fdv_amd import content
    kind, w, h, seed, kw = spec
    if kind == "objects":
        return content.scene_objects(w, h, seed, **kw)
    if kind == "pan":
        return content.scene_pan(w, h, seed, **kw)
    raise ValueError(kind)
