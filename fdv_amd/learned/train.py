"""Training of the learned motion proposer on the GPU, by the project's own
kernels (fdvgpu_nn_features / fdvgpu_nn_train_step / fdvgpu_nn_adam, include/
fdvgpu.h).  PyTorch only provides device memory and the stream.

Training tuples are generated by the GPU exhaustive search itself: every MB of
every P-frame of seeded synthetic sequences is coded by the exhaustive-table
wavefront (the reference's decision logic); the tuple is (the MB's proposer
input, the exhaustive 16x16 vector on the nearest reference as a position of
the half-res displacement grid, whether the MB was coded SKIP).

With --library N, the reference's scene library (bdrate.py --library) adds N
P-frames of every scene at every rd-bench QP, from frame --library-first on, so
that the frames the rate-distortion study codes (0 .. 29) are not trained on;
--holdout leaves whole scenes out.

    python -m fdv_amd.learned.train [--frames 2000] [--epochs 3] [--out fdv_amd/learned/proposer.f32]
    python -m fdv_amd.learned.train --library 10 --holdout chroma,detail [--out ...]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fdv_amd import content  # noqa: E402
from fdv_amd import gpu as G  # noqa: E402

NP = 289
NFEAT = NP * 2


def target_index(mvx, mvy):
    """Quarter-pel vectors -> positions of the 17x17 half-res displacement grid."""
    fx, fy = np.rint(np.asarray(mvx) / 4.0), np.rint(np.asarray(mvy) / 4.0)
    hx = np.clip(np.rint(fx / 2.0), -8, 8).astype(np.int64)
    hy = np.clip(np.rint(fy / 2.0), -8, 8).astype(np.int64)
    return ((hy + 8) * 17 + hx + 8).astype(np.int32)


def init_weights(seed: int = 0) -> np.ndarray:
    """He initialisation of the convolutions, small linear layer, zero biases
    (the layout of fdvgpu_nn.hip)."""
    rng = np.random.default_rng(seed)
    n = G.nn_nparam()
    W = np.zeros(n, dtype=np.float64)
    o = 0
    W[o:o + 576] = rng.normal(0, np.sqrt(2.0 / 18), 576); o += 576 + 32
    W[o:o + 9216] = rng.normal(0, np.sqrt(2.0 / 288), 9216); o += 9216 + 32
    W[o:o + 9216] = rng.normal(0, np.sqrt(2.0 / 288), 9216); o += 9216 + 32
    W[o:o + 64] = rng.normal(0, np.sqrt(1.0 / 32), 64)
    return W.astype(np.float32)


def make_scene(k: int, w: int, h: int):
    """Sequence k of the training set: config-1-style objects over blurred noise,
    alternating with panning scenes with grain."""
    rng = np.random.default_rng(7919 * k + 17)
    if k % 2 == 0:
        return content.scene_objects(w, h, 1000 + k, nobj=int(rng.integers(3, 9)), smin=16, smax=96,
                                     vmax=int(rng.integers(2, 13)), grain=float(rng.choice([0.0, 1.0, 2.0])))
    pan = (int(rng.integers(-6, 7)), int(rng.integers(-4, 5)))
    return content.scene_pan(w, h, 2000 + k, pan=pan, nobj=int(rng.integers(2, 8)), smin=16, smax=96,
                             vmax=int(rng.integers(2, 13)), grain=float(rng.choice([1.0, 2.0, 3.0])))


def generate(nseq: int, frames_per_seq: int, w: int, h: int, qps=(20, 24, 28, 32), first_seq: int = 0,
             log=print):
    """(features [N, 578] on the device, targets [N], SKIP labels [N]) as torch tensors."""
    items = [(make_scene(k, w, h), int(qps[k % len(qps)])) for k in range(first_seq, first_seq + nseq)]
    return generate_from(items, frames_per_seq, log)


def library_items(names, frames_per_seq: int, first: int, w: int, h: int, qps=(16, 22, 28, 34)):
    """(scene, qp) pairs of the reference's scene library: every scene at every
    QP, frames first .. first + frames_per_seq (frames 0 .. first - 1 are left to
    the rate-distortion study)."""
    from fdv_amd.learned import bdrate
    out = []
    for name, sc in bdrate.library_scenes(w, h, first + frames_per_seq + 1, names):
        sc.data = sc.data[first:]
        for qp in qps:
            out.append((sc, qp))
    return out


def generate_from(items, frames_per_seq: int, log=print):
    """Training tuples of (scene, qp) items, frames_per_seq P-frames each."""
    import torch
    L = G.lib()
    N = sum(frames_per_seq * (sc.w // 16) * (sc.h // 16) for sc, _ in items)
    feats = torch.empty((N, NFEAT), dtype=torch.float32, device="cuda")
    tgt = np.zeros(N, dtype=np.int32)
    skip = np.zeros(N, dtype=np.uint8)
    off = 0
    t0 = time.time()
    for k, (sc, qp) in enumerate(items):
        w, h = sc.w, sc.h
        nmb = (w // 16) * (h // 16)
        pos = torch.tensor([[(i % (w // 16)) * 16, (i // (w // 16)) * 16] for i in range(nmb)], dtype=torch.int32,
                           device="cuda")
        cw = torch.full((nmb,), w, dtype=torch.int32, device="cuda")
        an = G.Analyser([(sc.w, sc.h)], nsrc=1)
        an.put_ref(0, 0, sc.frame(0))
        for f in range(1, frames_per_seq + 1):
            an.put_source(0, 0, sc.frame(f))
            cy, _, ry, rs = an.planes(0, 0)
            curs = torch.full((nmb,), cy, dtype=torch.int64, device="cuda")
            refs = torch.full((nmb,), ry, dtype=torch.int64, device="cuda")
            rss = torch.full((nmb,), rs, dtype=torch.int32, device="cuda")
            qpt = torch.full((nmb,), qp, dtype=torch.int32, device="cuda")
            rc = L.fdvgpu_nn_features(curs.data_ptr(), cw.data_ptr(), refs.data_ptr(), rss.data_ptr(),
                                      pos.data_ptr(), qpt.data_ptr(), nmb, feats[off:off + nmb].data_ptr(), None)
            if rc != 0:
                raise G.GpuError("fdvgpu_nn_features failed")
            torch.cuda.synchronize()
            an.pframe([0], [qp], navail=[1 if f == 1 else 2], me_mode=G.ME_EXHAUSTIVE)
            an.fetch()
            mt = an.motion(0)
            dec = an.decisions(0)
            tgt[off:off + nmb] = target_index(mt["mv16"][:, 0, 0], mt["mv16"][:, 0, 1])
            skip[off:off + nmb] = (dec["mode"] == 0).astype(np.uint8)
            off += nmb
        an.close()
        if (k + 1) % 10 == 0:
            log("  data: %d/%d sequences, %d tuples, %.1f s" % (k + 1, len(items), off, time.time() - t0))
    return feats, torch.tensor(tgt, device="cuda"), torch.tensor(skip, device="cuda")


THRESHOLD = 0.7          # the proposer's confidence threshold (fdvgpu_api.hip)


def _metrics(l, n, nv, prefix=""):
    return {prefix + "vector_ce": float(l[0] / nv), prefix + "skip_ce": float(l[1] / n),
            prefix + "top1": float(l[2] / nv), prefix + "top8": float(l[3] / nv),
            prefix + "skip_rate": float(l[4] / n), prefix + "skip_precision": float(l[5] / max(l[4], 1)),
            prefix + "confident_rate": float(l[6] / nv), prefix + "confident_top8": float(l[7] / max(l[6], 1))}


def train(feats, tgt, skip, epochs: int = 3, batch: int = 512, lr: float = 2e-3, seed: int = 0, W0=None,
          vfeats=None, vtgt=None, vskip=None, wfp: float = 3.0, log=print):
    import torch
    L = G.lib()
    n = G.nn_nparam()
    W = torch.tensor(init_weights(seed) if W0 is None else W0, dtype=torch.float32, device="cuda")
    m1 = torch.zeros(n, dtype=torch.float32, device="cuda")
    m2 = torch.zeros_like(m1)
    grad = torch.zeros_like(m1)
    loss = torch.zeros(8, dtype=torch.float32, device="cuda")
    N = feats.shape[0]
    nv_all = max(1, int((tgt >= 0).sum().item()))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    steps_per_epoch = N // batch
    total = epochs * steps_per_epoch
    step = 0
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(N, device="cuda", generator=gen)
        loss.zero_()
        t0 = time.time()
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            f = feats[idx].contiguous()
            t = tgt[idx].contiguous()
            s = skip[idx].contiguous()
            if L.fdvgpu_nn_train_step(W.data_ptr(), f.data_ptr(), t.data_ptr(), s.data_ptr(), batch, wfp, THRESHOLD,
                                      grad.data_ptr(), loss.data_ptr(), min(batch, 512), None) != 0:
                raise G.GpuError("fdvgpu_nn_train_step failed")
            step += 1
            cur_lr = lr * 0.5 * (1.0 + np.cos(np.pi * step / total))
            if L.fdvgpu_nn_adam(W.data_ptr(), grad.data_ptr(), m1.data_ptr(), m2.data_ptr(), n, float(cur_lr), step,
                                batch, None) != 0:
                raise G.GpuError("fdvgpu_nn_adam failed")
        torch.cuda.synchronize()
        l = loss.cpu().numpy()
        frac = steps_per_epoch * batch / N
        rec = {"epoch": ep + 1, "seconds": round(time.time() - t0, 2)}
        rec.update(_metrics(l, steps_per_epoch * batch, nv_all * frac))
        if vfeats is not None:
            rec.update(evaluate(W, vfeats, vtgt, vskip, wfp=wfp))
        hist.append(rec)
        log("  epoch %d: %s" % (ep + 1, json.dumps(rec)))
    return W.cpu().numpy(), hist


def evaluate(W, feats, tgt, skip, batch: int = 4096, wfp: float = 3.0):
    """Validation losses and hit rates (a training step with the gradient discarded)."""
    import torch
    L = G.lib()
    n = G.nn_nparam()
    grad = torch.zeros(n, dtype=torch.float32, device="cuda")
    loss = torch.zeros(8, dtype=torch.float32, device="cuda")
    N = feats.shape[0]
    for b in range(0, N, batch):
        m = min(batch, N - b)
        L.fdvgpu_nn_train_step(W.data_ptr(), feats[b:b + m].data_ptr(), tgt[b:b + m].data_ptr(),
                               skip[b:b + m].data_ptr(), m, wfp, THRESHOLD, grad.data_ptr(), loss.data_ptr(),
                               min(m, 512), None)
    torch.cuda.synchronize()
    nv = max(1, int((tgt >= 0).sum().item()))
    return _metrics(loss.cpu().numpy(), N, nv, "val_")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=2000, help="training P-frames")
    ap.add_argument("--frames-per-seq", type=int, default=20)
    ap.add_argument("--size", default="640x368")
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--wfp", type=float, default=3.0, help="weight of false SKIP positives in the loss")
    ap.add_argument("--out", default=G.PROPOSER_WEIGHTS)
    ap.add_argument("--library", type=int, default=0,
                    help="P-frames per scene of the reference's scene library to add (0: none)")
    ap.add_argument("--library-first", type=int, default=30, help="first library frame used for training")
    ap.add_argument("--library-size", default="1280x720")
    ap.add_argument("--holdout", default="", help="comma-separated library scenes left out of training")
    args = ap.parse_args()
    import torch
    w, h = (int(x) for x in args.size.split("x"))
    nseq = max(1, args.frames // args.frames_per_seq)
    print("generating %d sequences x %d P-frames at %dx%d" % (nseq, args.frames_per_seq, w, h), flush=True)
    feats, tgt, skip = generate(nseq, args.frames_per_seq, w, h, log=lambda m: print(m, flush=True))
    lib_names = []
    if args.library > 0:
        from fdv_amd.learned import bdrate
        import glob
        hold = set(x for x in args.holdout.split(",") if x)
        lib_names = [os.path.basename(p)[:-4] for p in sorted(glob.glob(os.path.join(bdrate.SCENE_DIR, "*.scn")))
                     if os.path.basename(p)[:-4] not in hold]
        lw, lh = (int(x) for x in args.library_size.split("x"))
        items = library_items(lib_names, args.library, args.library_first, lw, lh)
        print("generating %d library sequences x %d P-frames at %dx%d" % (len(items), args.library, lw, lh),
              flush=True)
        lf, lt, ls = generate_from(items, args.library, log=lambda m: print(m, flush=True))
        feats, tgt, skip = torch.cat([feats, lf]), torch.cat([tgt, lt]), torch.cat([skip, ls])
        del lf, lt, ls
    vf, vt, vs = generate(8, 10, w, h, first_seq=100000, log=lambda m: print(m, flush=True))
    print("training on %d tuples, validating on %d" % (feats.shape[0], vf.shape[0]), flush=True)
    W, hist = train(feats, tgt, skip, epochs=args.epochs, batch=args.batch, lr=args.lr, vfeats=vf, vtgt=vt,
                    vskip=vs, wfp=args.wfp, log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    W.astype("<f4").tofile(args.out)
    meta = {"tuples": int(feats.shape[0]), "frames": nseq * args.frames_per_seq, "size": [w, h],
            "qps": [20, 24, 28, 32], "epochs": args.epochs, "batch": args.batch, "lr": args.lr, "wfp": args.wfp,
            "library": {"scenes": lib_names, "frames_per_scene": args.library, "first_frame": args.library_first,
                        "size": args.library_size, "qps": [16, 22, 28, 34]} if lib_names else None,
            "history": hist}
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
