"""Deterministic test planes for the intra (key-frame) codec: each kind is
chosen to reach a different part of it -- flat areas that code as whole
64x64 leaves, texture that splits down to 4x4, gradients for the plane
predictor, noise, and periodic texture whose streams are best coded with a
table per stream."""
import numpy as np

KINDS = ("mixed", "noise", "checker", "ramp", "flat", "sine", "halfnoise")


def intra_plane(kind: str, w: int, h: int, seed: int) -> np.ndarray:
    r = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    if kind == "mixed":
        p = 128 + 60 * np.sin(x / (7 + seed % 5)) * np.cos(y / 11) + 0.3 * x + r.normal(0, 6 + seed % 3 * 8, (h, w))
        band = (x // 16 + y // 16) % 3 == 0
        p[band] += 40 * ((x + y) % 8 < 4)[band]
        if seed % 2:
            p[h // 3: h // 2, :] = 30
    elif kind == "noise":
        p = r.integers(0, 256, (h, w))
    elif kind == "checker":
        p = ((x // 2 + y // 2) % 2) * 255
    elif kind == "ramp":
        p = (2 * x + y + seed) % 256
    elif kind == "flat":
        p = np.full((h, w), 40 + seed % 150)
    elif kind == "sine":
        p = 128 + 100 * np.sin(x * 0.9) * np.sin(y * 0.7)
    else:
        p = np.where(y < h // 2, r.integers(0, 256, (h, w)), 90)
    return np.ascontiguousarray(np.clip(p, 0, 255).astype(np.uint8))


# (kind, w, h, qp, seed): sizes with and without whole 64x64 units, QPs on both
# sides of the screen's threshold (12), and containers of all three kinds
CASES = [
    ("mixed", 64, 64, 24, 1), ("mixed", 72, 40, 36, 2), ("mixed", 128, 96, 4, 3),
    ("mixed", 100, 36, 12, 4), ("mixed", 192, 68, 11, 5), ("mixed", 260, 132, 30, 6),
    ("mixed", 4, 4, 20, 7), ("mixed", 8, 8, 11, 8), ("noise", 96, 116, 51, 9),
    ("checker", 80, 48, 26, 10), ("ramp", 136, 72, 18, 11), ("flat", 124, 112, 13, 12),
    ("sine", 192, 120, 1, 13), ("sine", 168, 144, 18, 14), ("halfnoise", 64, 40, 40, 15),
    ("mixed", 1280, 96, 24, 16),
]
