"""Rate-distortion comparison of two search modes (BD-rate), as the reference's
tools/rd-bench.py:75-113 defines it:
  - drop dominated operating points (keep a point only if it has more PSNR than
    every cheaper one);
  - fit log10(rate) as a polynomial of degree min(3, n - 1) in PSNR;
  - integrate both fits over the overlapping PSNR range;
  - BD-rate = 10^(mean difference) - 1, in percent (negative: the new mode
    needs fewer bits for equal quality).

evaluate() codes the same sequences with two modes of the drop-in encoder at
several QPs (payload bytes of the P-frames; PSNR-Y of their reconstruction
against the source) and returns the BD-rate of the second against the first.

Content: either two seeded sequences of fdv_amd/content.py, or (--library) the
reference's scene library (scenes/*.scn, rendered by its include/fdv_scene.h
through scene_render.c; `make -C fdv_amd/learned scenes` builds both into
_scenes/) at rd-bench.py's QPs (16, 22, 28, 34: tools/rd-bench.py:19), with its
per-scene summary (mean, median, best, worst: tools/rd-bench.py:116-138).

    python -m fdv_amd.learned.bdrate [--frames 60] [--qps 20,24,28,32]
    python -m fdv_amd.learned.bdrate --library [--frames 30] [--size 1280x720]
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)


def frontier(pts):
    out, best = [], float("-inf")
    for r, d in sorted(pts):
        if d > best:
            out.append((r, d))
            best = d
    return out


def bd_rate(ref, new):
    ref, new = frontier(ref), frontier(new)
    if len(ref) < 3 or len(new) < 3:
        return None

    def fit(pts):
        d = np.array([p[1] for p in pts])
        r = np.log10([p[0] for p in pts])
        return np.polyfit(d, r, min(3, len(pts) - 1)), d.min(), d.max()

    pr, lr, hr = fit(ref)
    pn, ln, hn = fit(new)
    lo, hi = max(lr, ln), min(hr, hn)
    if hi <= lo:
        return None
    ir = np.polyval(np.polyint(pr), [lo, hi])
    inw = np.polyval(np.polyint(pn), [lo, hi])
    return float((10 ** (((inw[1] - inw[0]) - (ir[1] - ir[0])) / (hi - lo)) - 1) * 100)


def psnr_y(rec, src, n):
    e = np.mean((rec[:n].astype(np.float64) - src[:n].astype(np.float64)) ** 2)
    return 99.0 if e == 0 else 10 * np.log10(255.0 ** 2 / e)


def bands_of(h, band_mbrows=6):
    bh = band_mbrows * 16
    return [(y, min(bh, h - y)) for y in range(0, h, bh)]


def band_frame(frame, w, h, y0, bh):
    ys, cw, ch = w * h, w // 2, h // 2
    Y = frame[:ys].reshape(h, w)
    U = frame[ys:ys + cw * ch].reshape(ch, cw)
    V = frame[ys + cw * ch:].reshape(ch, cw)
    return np.concatenate([Y[y0:y0 + bh].ravel(), U[y0 // 2:(y0 + bh) // 2].ravel(), V[y0 // 2:(y0 + bh) // 2].ravel()])


def code_sequence(frames, w, h, qp, mode, weights=None, thresholds=None):
    """Bytes and mean PSNR-Y of P-frames 1.. of `frames` (frame 0 is installed as
    the key frame's reconstruction), coded as 6-MB-row bands in one batch."""
    from fdv_amd import gpu as G
    layout = bands_of(h)
    enc = G.Encoder([(w, bh) for _, bh in layout], me_mode=mode)
    if mode == G.ME_LEARNED:
        enc.set_proposer(weights)
        if thresholds is not None:
            enc.set_proposer_thresholds(*thresholds)
    for b, (y0, bh) in enumerate(layout):
        enc.set_ref(b, 0, band_frame(frames[0], w, h, y0, bh))
    tcs = [G.TabCache() for _ in layout]
    total, ps = 0, []
    for f in range(1, len(frames)):
        fb = [band_frame(frames[f], w, h, y0, bh) for y0, bh in layout]
        pays, recs = enc.pframe_batch(fb, [1 if f == 1 else 2] * len(layout), [qp] * len(layout), tcs)
        total += sum(len(p) for p in pays)
        se = sum(float(np.sum((r[:w * bh].astype(np.float64) - s[:w * bh]) ** 2))
                 for r, s, (_, bh) in zip(recs, fb, layout))
        mse = se / (w * h)
        ps.append(99.0 if mse == 0 else 10 * np.log10(255.0 ** 2 / mse))
    enc.close()
    return total, float(np.mean(ps))


def evaluate(scenes, qps=(20, 24, 28, 32), nframes=60, mode_a=None, mode_b=None, weights=None, log=print,
             thresholds=None, cache=None):
    """`cache` (a dict, optional) keeps each (scene, QP, mode) result, so that one
    first mode can be compared with several second ones without recoding it."""
    from fdv_amd import gpu as G
    mode_a = G.ME_EXHAUSTIVE if mode_a is None else mode_a
    mode_b = G.ME_LEARNED if mode_b is None else mode_b
    cache = {} if cache is None else cache

    def coded(name, frames, sc, qp, mode):
        key = (name, qp, mode)
        if key not in cache:
            cache[key] = code_sequence(frames, sc.w, sc.h, qp, mode, weights, thresholds)
        return cache[key]

    out = {}
    for name, sc in scenes:
        frames = [sc.frame(i) for i in range(nframes)]
        pa, pb = [], []
        for qp in qps:
            ba, da = coded(name, frames, sc, qp, mode_a)
            bb, db = coded(name, frames, sc, qp, mode_b)
            pa.append((ba, da))
            pb.append((bb, db))
            log("  %s qp %d: A %d B %.3f dB | B %d B %.3f dB" % (name, qp, ba, da, bb, db))
        out[name] = {"a": pa, "b": pb, "bd_rate": bd_rate(pa, pb)}
    return out


SCENE_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_scenes")


class LibraryScene:
    """One scene of the reference's library, rendered once to a temporary file."""

    def __init__(self, path, w, h, nframes):
        import subprocess
        import tempfile
        exe = os.path.join(SCENE_DIR, "scene_render")
        if not os.path.exists(exe):
            raise RuntimeError("scene library not built: make -C fdv_amd/learned scenes")
        self.w, self.h = w, h
        fs = w * h * 3 // 2
        with tempfile.NamedTemporaryFile(suffix=".yuv") as t:
            subprocess.run([exe, path, str(w), str(h), str(nframes), t.name], check=True)
            self.data = np.fromfile(t.name, dtype=np.uint8).reshape(nframes, fs)

    def frame(self, i):
        return self.data[i]


def library_scenes(w, h, nframes, names=None):
    """(name, scene) for every .scn of the library, in name order."""
    import glob
    out = []
    for path in sorted(glob.glob(os.path.join(SCENE_DIR, "*.scn"))):
        name = os.path.basename(path)[:-4]
        if names and name not in names:
            continue
        out.append((name, LibraryScene(path, w, h, nframes)))
    return out


def summarise(res):
    """rd-bench.py's per-scene summary (tools/rd-bench.py:116-138)."""
    rows = sorted(((k, v["bd_rate"]) for k, v in res.items() if v["bd_rate"] is not None), key=lambda r: r[1])
    if not rows:
        return None
    v = [r[1] for r in rows]
    return {"scenes": len(rows), "mean": float(np.mean(v)), "median": float(np.median(v)),
            "best": list(rows[0]), "worst": list(rows[-1]),
            "improved": sum(1 for x in v if x < 0), "regressed": sum(1 for x in v if x > 0)}


def main():
    from fdv_amd import content
    from fdv_amd import gpu as G
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=None, help="frames per sequence (60; 30 with --library)")
    ap.add_argument("--qps", default=None, help="QPs (20,24,28,32; rd-bench's 16,22,28,34 with --library)")
    ap.add_argument("--library", action="store_true", help="the reference's scene library as content")
    ap.add_argument("--size", default="1280x720", help="library render size")
    ap.add_argument("--scenes", default=None, help="comma-separated subset of the library")
    ap.add_argument("--b", default="learned",
                    help="second mode(s), comma-separated: learned, parallel, exhaustive, diamond")
    ap.add_argument("--a", default="exhaustive", choices=["learned", "parallel", "exhaustive", "diamond"])
    ap.add_argument("--out", default=None, help="write the points and BD-rates as JSON")
    ap.add_argument("--conf", type=float, default=0.7, help="proposer confidence threshold")
    ap.add_argument("--skip", type=float, default=2.0, help="proposer SKIP threshold (above 1: no pruning)")
    ap.add_argument("--weights", default=None, help="proposer weights file (default: the shipped ones)")
    args = ap.parse_args()
    weights = np.fromfile(args.weights, dtype="<f4") if args.weights else None
    modes = {"learned": G.ME_LEARNED, "parallel": G.ME_PARALLEL, "exhaustive": G.ME_EXHAUSTIVE,
             "diamond": G.ME_DIAMOND}
    if args.library:
        frames = args.frames or 30
        qps = args.qps or "16,22,28,34"
        w, h = (int(x) for x in args.size.split("x"))
        scenes = library_scenes(w, h, frames, args.scenes.split(",") if args.scenes else None)
    else:
        frames = args.frames or 60
        qps = args.qps or "20,24,28,32"
        scenes = [("pan720", content.scene_pan(1280, 720, 2)),
                  ("objects720", content.scene_objects(1280, 720, 3, nobj=12, smin=32, smax=160, vmax=12,
                                                       grain=2.0))]
    cache, recs = {}, {}
    for bname in args.b.split(","):
        print("== %s vs %s" % (bname, args.a), flush=True)
        res = evaluate(scenes, [int(q) for q in qps.split(",")], frames, modes[args.a], modes[bname], weights=weights,
                       log=lambda m: print(m, flush=True), thresholds=(args.conf, args.skip), cache=cache)
        summary = {k: v["bd_rate"] for k, v in res.items()}
        print(json.dumps(summary), flush=True)
        stats = summarise(res)
        if stats:
            print("%s vs %s: mean %+.2f%%, median %+.2f%%, best %s %+.2f%%, worst %s %+.2f%%, improved %d/%d" % (
                bname, args.a, stats["mean"], stats["median"], stats["best"][0], stats["best"][1],
                stats["worst"][0], stats["worst"][1], stats["improved"], stats["scenes"]), flush=True)
        recs[bname] = {"a": args.a, "b": bname, "frames": frames, "qps": qps,
                       "content": ("reference scene library at " + args.size) if args.library
                       else "content.py seeds",
                       "thresholds": {"conf": args.conf, "skip": args.skip},
                       "bd_rate_percent": summary, "summary": stats,
                       "points": {k: {"a": v["a"], "b": v["b"]} for k, v in res.items()}}
    if args.out:
        with open(args.out, "w") as f:
            json.dump(recs[args.b] if "," not in args.b else recs, f, indent=1)


if __name__ == "__main__":
    main()
