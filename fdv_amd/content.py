"""Seeded synthetic I420 content for the encode-analysis workloads.

The reference renders content from scene files (include/fdv_scene.h); the
workloads this path is measured on are defined by a seed instead, so this module
is the project's own deterministic generator.  Every frame is a pure function of
(seed, frame index): textures are built once from integer noise blurred with a
fixed integer Gaussian kernel, objects move at integer velocities, and grain is
drawn per frame from a generator seeded by (seed, frame).

Frames are tight I420 byte arrays (w*h luma, then (w/2)*(h/2) U and V).  Heights
that are not a multiple of 16 (180, 360, 1080) are padded to the next multiple
by replicating the last row, because the codec requires 16-aligned dimensions
(fdv.h:6743).
"""
from __future__ import annotations

import numpy as np

__all__ = ["pad16", "Scene", "scene_objects", "scene_pan", "scene_blobs", "i420_size"]


def pad16(n: int) -> int:
    return (n + 15) & ~15


def i420_size(w: int, h: int) -> int:
    return w * h + 2 * (w // 2) * (h // 2)


def _gauss_taps(sigma: float) -> np.ndarray:
    r = int(3 * sigma)
    k = np.arange(-r, r + 1, dtype=np.float64)
    t = np.rint(4096.0 * np.exp(-(k * k) / (2.0 * sigma * sigma))).astype(np.int64)
    return t


def _blur(img: np.ndarray, sigma: float) -> np.ndarray:
    """Separable integer Gaussian with edge replication; returns uint8."""
    t = _gauss_taps(sigma)
    r = len(t) // 2
    s = int(t.sum())
    a = img.astype(np.int64)
    p = np.pad(a, ((0, 0), (r, r)), mode="edge")
    acc = np.zeros_like(a)
    for i, tv in enumerate(t):
        acc += tv * p[:, i:i + a.shape[1]]
    a = (acc + s // 2) // s
    p = np.pad(a, ((r, r), (0, 0)), mode="edge")
    acc = np.zeros_like(a)
    for i, tv in enumerate(t):
        acc += tv * p[i:i + a.shape[0], :]
    a = (acc + s // 2) // s
    return np.clip(a, 0, 255).astype(np.uint8)


def _noise_texture(rng: np.random.Generator, h: int, w: int, sigma: float = 4.0,
                   lo: int = 16, hi: int = 235) -> np.ndarray:
    n = rng.integers(lo, hi + 1, size=(h, w), dtype=np.int64)
    return _blur(n, sigma)


class Scene:
    """A deterministic sequence: frame(i) -> tight I420 bytes (padded height)."""

    def __init__(self, w: int, h: int, render):
        self.w = w
        self.h_src = h
        self.h = pad16(h)
        self._render = render

    @property
    def frame_bytes(self) -> int:
        return i420_size(self.w, self.h)

    def planes(self, i: int):
        y, u, v = self._render(i)
        if self.h != self.h_src:
            y = np.concatenate([y, np.repeat(y[-1:], self.h - self.h_src, axis=0)], 0)
            ch = self.h // 2
            u = np.concatenate([u, np.repeat(u[-1:], ch - u.shape[0], axis=0)], 0)
            v = np.concatenate([v, np.repeat(v[-1:], ch - v.shape[0], axis=0)], 0)
        return (np.ascontiguousarray(y, dtype=np.uint8), np.ascontiguousarray(u, dtype=np.uint8),
                np.ascontiguousarray(v, dtype=np.uint8))

    def frame(self, i: int) -> np.ndarray:
        y, u, v = self.planes(i)
        return np.concatenate([y.ravel(), u.ravel(), v.ravel()])


def _chroma_base(w: int, h: int):
    cw, ch = w // 2, (h + 1) // 2
    x = np.arange(cw, dtype=np.int64)
    gu = 96 + (64 * x) // max(cw - 1, 1)
    gv = 160 - (64 * x) // max(cw - 1, 1)
    return np.broadcast_to(gu, (ch, cw)).copy(), np.broadcast_to(gv, (ch, cw)).copy()


def _place(dst: np.ndarray, tex: np.ndarray, x0: int, y0: int):
    """Paste tex at (x0, y0), clipped to dst."""
    H, W = dst.shape
    th, tw = tex.shape
    xa, ya = max(x0, 0), max(y0, 0)
    xb, yb = min(x0 + tw, W), min(y0 + th, H)
    if xa >= xb or ya >= yb:
        return
    dst[ya:yb, xa:xb] = tex[ya - y0:yb - y0, xa - x0:xb - x0]


def _objects(rng, w, h, n, smin, smax, vmax):
    objs = []
    for _ in range(n):
        ow = int(rng.integers(smin, smax + 1))
        oh = int(rng.integers(smin, smax + 1))
        objs.append(dict(
            w=ow, h=oh,
            x=int(rng.integers(-ow // 2, w - ow // 2)), y=int(rng.integers(-oh // 2, h - oh // 2)),
            vx=int(rng.integers(-vmax, vmax + 1)), vy=int(rng.integers(-vmax, vmax + 1)),
            tex=_noise_texture(rng, oh, ow),
            du=int(rng.integers(-24, 25)), dv=int(rng.integers(-24, 25)),
        ))
    return objs


def _wrap(p: int, size: int, extent: int) -> int:
    """Objects re-enter from the other side so they stay on screen."""
    span = size + extent
    return ((p + extent) % span) - extent


def scene_objects(w: int, h: int, seed: int, nobj: int = 4, smin: int = 24, smax: int = 64,
                  vmax: int = 8, grain: float = 0.0) -> Scene:
    """Blurred-noise background with textured rectangles at constant integer
    velocities; chroma is a horizontal gradient plus per-object offsets."""
    rng = np.random.default_rng(seed)
    bg = _noise_texture(rng, h, w)
    objs = _objects(rng, w, h, nobj, smin, smax, vmax)
    gu, gv = _chroma_base(w, h)

    def render(i):
        y = bg.copy()
        u = gu.copy()
        v = gv.copy()
        for o in objs:
            x0 = _wrap(o["x"] + o["vx"] * i, w, o["w"])
            y0 = _wrap(o["y"] + o["vy"] * i, h, o["h"])
            _place(y, o["tex"], x0, y0)
            cu = np.full((max(o["h"] // 2, 1), max(o["w"] // 2, 1)), 0, dtype=np.int64)
            _place(u, np.clip(cu + 128 + o["du"], 0, 255), x0 // 2, y0 // 2)
            _place(v, np.clip(cu + 128 + o["dv"], 0, 255), x0 // 2, y0 // 2)
        if grain > 0:
            g = np.random.default_rng([seed, i, 7]).normal(0.0, grain, size=y.shape)
            y = np.clip(y.astype(np.int64) + np.rint(g).astype(np.int64), 0, 255)
        return y, np.clip(u, 0, 255), np.clip(v, 0, 255)

    return Scene(w, h, render)


def scene_pan(w: int, h: int, seed: int, pan=(3, -1), nobj: int = 12, smin: int = 32,
              smax: int = 160, vmax: int = 16, grain: float = 3.0) -> Scene:
    """A global pan over a 2x-oversized blurred-noise texture, textured rectangles
    with their own velocities, and fresh Gaussian grain every frame."""
    rng = np.random.default_rng(seed)
    big = _noise_texture(rng, 2 * h, 2 * w)
    objs = _objects(rng, w, h, nobj, smin, smax, vmax)
    gu, gv = _chroma_base(w, h)

    def render(i):
        ox = (w // 2 + pan[0] * i) % w
        oy = (h // 2 + pan[1] * i) % h
        y = big[oy:oy + h, ox:ox + w].copy()
        u = gu.copy()
        v = gv.copy()
        for o in objs:
            x0 = _wrap(o["x"] + o["vx"] * i, w, o["w"])
            y0 = _wrap(o["y"] + o["vy"] * i, h, o["h"])
            _place(y, o["tex"], x0, y0)
            c = np.zeros((max(o["h"] // 2, 1), max(o["w"] // 2, 1)), dtype=np.int64)
            _place(u, np.clip(c + 128 + o["du"], 0, 255), x0 // 2, y0 // 2)
            _place(v, np.clip(c + 128 + o["dv"], 0, 255), x0 // 2, y0 // 2)
        if grain > 0:
            g = np.random.default_rng([seed, i, 11]).normal(0.0, grain, size=y.shape)
            y = np.clip(y.astype(np.int64) + np.rint(g).astype(np.int64), 0, 255)
        return y, np.clip(u, 0, 255), np.clip(v, 0, 255)

    return Scene(w, h, render)


def scene_blobs(w: int, h: int, seed: int, nblob: int = 500, vmax: int = 48, cut: int = 40,
                grain: float = 5.0) -> Scene:
    """Many small coloured blobs at high velocity over a blurred-noise background,
    a hard cut to a new background every `cut` frames, and grain."""
    rng = np.random.default_rng(seed)
    sz = rng.integers(4, 13, size=nblob)
    px = rng.integers(0, w, size=nblob)
    py = rng.integers(0, h, size=nblob)
    vx = rng.integers(-vmax, vmax + 1, size=nblob)
    vy = rng.integers(-vmax, vmax + 1, size=nblob)
    col = rng.integers(16, 236, size=(nblob, 3))
    gu, gv = _chroma_base(w, h)
    bgs = {}

    def background(k):
        if k not in bgs:
            bgs.clear()
            bgs[k] = _noise_texture(np.random.default_rng([seed, k, 3]), h, w)
        return bgs[k]

    def render(i):
        y = background(i // cut).copy()
        u = gu.copy()
        v = gv.copy()
        for b in range(nblob):
            s = int(sz[b])
            x0 = int((px[b] + vx[b] * i) % (w + s)) - s
            y0 = int((py[b] + vy[b] * i) % (h + s)) - s
            _place(y, np.full((s, s), col[b, 0], dtype=np.int64), x0, y0)
            cs = max(s // 2, 1)
            _place(u, np.full((cs, cs), col[b, 1], dtype=np.int64), x0 // 2, y0 // 2)
            _place(v, np.full((cs, cs), col[b, 2], dtype=np.int64), x0 // 2, y0 // 2)
        if grain > 0:
            g = np.random.default_rng([seed, i, 13]).normal(0.0, grain, size=y.shape)
            y = np.clip(y.astype(np.int64) + np.rint(g).astype(np.int64), 0, 255)
        return y, np.clip(u, 0, 255), np.clip(v, 0, 255)

    return Scene(w, h, render)
