"""Encode-analysis throughput on one MI355X.

Workload (the configuration the headline metric is quoted on): 1280x720 4:2:0,
QP 24, 8 tiles -- the reference's tile-parallel video with 6 MB rows per band
(fdv_vtile_encode, fdv.h:7161; 7 bands of 1280x96 and one of 1280x48) -- on
seeded synthetic content (global pan over a blurred-noise texture, 12 textured
rectangles, Gaussian grain sigma 3).  Frame 0 is the key frame: its source is
installed as the reference (I-frame coding is outside this path); every step
is one P-frame of all 8 bands in one batched device pass: motion search, mode
decision, residual coding, reconstruction, deblocking and serialisation of the
16 symbol streams, plus their copy back to the host.  Source frames are resident
in HBM before timing starts.  Steps are queued two deep (a step's
serialisation and stream staging run beside the next step's motion tables, and
each step's streams are collected once the next is queued), for every mode
compared; --sync times them one at a time.

--mode picks the search: learned (default: the learned proposer with the
exhaustive search below its confidence threshold, block-parallel decisions),
exhaustive, parallel (exhaustive search, block-parallel decisions) or diamond
(the reference's search and decision wavefront, bit-exact with pframe_encode).
The exhaustive mode is timed on the same frames for the learned-vs-exhaustive
comparison; launch shapes are autotuned on the first frame.

Prints one JSON line (see the keys below).  --gpus above 1 is refused: this
path runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "encode-analysis Mpix/s at 720p/1080p, learned vs exhaustive, at ≤1.5% BD-rate"
HBM_PEAK_GBS = 8000.0


def bands(h: int, band_mbrows: int):
    bh = band_mbrows * 16
    out, y = [], 0
    while y < h:
        out.append((y, min(bh, h - y)))
        y += bh
    return out


def band_frames(frame: np.ndarray, w: int, h: int, layout):
    """Split a tight I420 frame into the tight I420 frames of its bands."""
    ys = w * h
    cw, ch = w // 2, h // 2
    Y = frame[:ys].reshape(h, w)
    U = frame[ys:ys + cw * ch].reshape(ch, cw)
    V = frame[ys + cw * ch:].reshape(ch, cw)
    out = []
    for y0, bh in layout:
        out.append(np.concatenate([Y[y0:y0 + bh].ravel(), U[y0 // 2:(y0 + bh) // 2].ravel(),
                                   V[y0 // 2:(y0 + bh) // 2].ravel()]))
    return out


def cpu_baseline(frames, w, layout, qp, seconds=12.0):
    """The reference P-frame encoder (built from its own sources, when present)
    or else the oracle port, band-parallel like the reference's tiled encoder
    (one thread per band, fdv_vtile_encode at fdv.h:7161), on the frames of the
    workload, cycled until about `seconds` of wall time have passed."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import orclib as L
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "Mpix/s", "cores": 0, "kind": "port", "sample": "unavailable: %s" % e}
    use_ref = L.ref() is not None
    kind = "reference" if use_ref else "port"
    h = sum(b for _, b in layout)
    per = [band_frames(f, w, h, layout) for f in frames]
    nb = len(layout)

    def band_seq(b, nfr):
        tc = L.TabCache()
        near, far = per[0][b], None
        for i in range(nfr):
            f = 1 + i % (len(frames) - 1)
            navail = 1 if i == 0 else 2
            if use_ref:
                _, rec = L.ref_pframe(per[f][b], near, far, w, layout[b][1], navail, qp, tc)
            else:
                _, rec, _ = L.oracle_pframe(per[f][b], near, far, w, layout[b][1], navail, qp, tc)
            far, near = near, rec

    # size the sample from a short single-frame probe, then time the full run
    t0 = time.perf_counter()
    with ThreadPoolExecutor(nb) as ex:
        list(ex.map(lambda b: band_seq(b, 1), range(nb)))
    probe = time.perf_counter() - t0
    nfr = max(2, int(seconds / max(probe, 1e-3)))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(nb) as ex:
        list(ex.map(lambda b: band_seq(b, nfr), range(nb)))
    dt = time.perf_counter() - t0
    px = nfr * w * h
    return {"value": round(px / dt / 1e6, 3), "unit": "Mpix/s", "cores": nb, "kind": kind,
            "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
            "sample": "%d P-frames of the same 1280x720 8-band workload (frames cycled), "
                      "%d threads, one per band (%.1f s)" % (nfr, nb, dt),
            "includes": "the reference's whole P-frame encode (pframe_encode), entropy coding of the payload "
                        "included; the GPU value excludes entropy coding (end_to_end includes it)"}


def bench_keyframe(args, content, G):
    """Key frames (the device intra quadtree, fdvamd_image_encode_batch): frame
    0 of the headline 1280x720 scene as 8 bands of 6 MB rows (all 24 planes in
    one device pass) and as one whole frame, QP 24, timed from host buffers to
    containers (uploads, device decisions, host serialisation and entropy
    coding); beside it the reference's key-frame coder (iframe_encode) on the
    same planes, one thread per band."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    W, H, QP = 1280, 720, 24
    f = content.scene_pan(W, H, seed=2).frame(0)
    Y = f[:W * H].reshape(H, W)
    U = f[W * H:W * H * 5 // 4].reshape(H // 2, W // 2)
    V = f[W * H * 5 // 4:].reshape(H // 2, W // 2)
    bands_ = [(b * 96, 96) for b in range(8)]
    tiled = []
    for y0, bh in bands_:
        tiled += [np.ascontiguousarray(Y[y0:y0 + bh]), np.ascontiguousarray(U[y0 // 2:(y0 + bh) // 2]),
                  np.ascontiguousarray(V[y0 // 2:(y0 + bh) // 2])]
    whole = [np.ascontiguousarray(Y), np.ascontiguousarray(U), np.ascontiguousarray(V)]
    res = {}
    for name, planes in (("tiled", tiled), ("whole", whole)):
        for _ in range(max(1, args.warmup)):
            G.image_encode_batch(planes, [QP] * len(planes))
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = G.image_encode_batch(planes, [QP] * len(planes))
        res[name] = (time.perf_counter() - t0) / args.steps
        res[name + "_bytes"] = sum(len(o) for o, _ in out)
    cpu = None
    if not args.no_cpu:
        try:
            import orclib as L
            lib, fn, kind = (L.ref(), "fdvref_iframe_encode", "reference") if L.ref() is not None else \
                (L.oracle(), "orc_iframe_encode", "port")

            def band(b):
                y, u, v = tiled[3 * b], tiled[3 * b + 1], tiled[3 * b + 2]
                cap = y.size * 4 + 65536
                o = np.zeros(cap, np.uint8)
                r = [np.zeros_like(y), np.zeros_like(u), np.zeros_like(v)]
                getattr(lib, fn)(y.ctypes.data, u.ctypes.data, v.ctypes.data, W, y.shape[0], QP, o.ctypes.data, cap,
                                 r[0].ctypes.data, r[1].ctypes.data, r[2].ctypes.data)
            reps = 3
            t0 = time.perf_counter()
            for _ in range(reps):
                th = [threading.Thread(target=band, args=(b,)) for b in range(8)]
                [t.start() for t in th]
                [t.join() for t in th]
            dt = (time.perf_counter() - t0) / reps
            cpu = {"value": round(W * H / dt / 1e6, 3), "unit": "Mpix/s", "ms_per_key_frame": round(dt * 1e3, 2),
                   "cores": 8, "kind": kind, "nproc": os.cpu_count(), "cpu_model": _cpu_model(),
                   "sample": "%d key frames of the same 8 bands, one thread per band" % reps}
        except Exception as e:  # pragma: no cover
            cpu = {"value": None, "sample": "unavailable: %s" % e}
    print(json.dumps({
        "metric": "key-frame encode Mpix/s at 720p (8 bands)", "value": round(W * H / res["tiled"] / 1e6, 3),
        "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(res["tiled"] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8", "data": "synthetic (seeded pan + objects + grain), frame 0",
        "config": {"workload": "1280x720 key frame, QP 24, 8 bands of 6 MB rows: 24 planes in one device pass",
                   "container_bytes": res["tiled_bytes"]},
        "whole_frame": {"ms_per_key_frame": round(res["whole"] * 1e3, 3),
                        "value": round(W * H / res["whole"] / 1e6, 3), "unit": "Mpix/s",
                        "container_bytes": res["whole_bytes"]},
        "cpu_baseline": cpu}))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def profiled_traffic(mode: str):
    """HBM bytes per launch of the dominant kernel (diamond / exhaustive: the
    k_analyse wavefront; learned / parallel: the decision stage's three kernels,
    k_pcosts, k_pxrep and k_pdecide)
    from the committed rocprofv3 counter summary of the 720p workload
    (FETCH_SIZE and WRITE_SIZE, each calibrated against a plain copy), or None
    when no summary exists for this mode."""
    name = ("decisions_720p_%s.md" if mode in ("learned", "parallel") else "k_analyse_720p_%s.md") % mode
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def _job_720p(content, nframes):
    """The headline workload: one 1280x720 stream (pan over a blurred-noise
    texture, 12 textured rectangles, grain sigma 3), QP 24, as 8 bands of 6 MB
    rows coded in one batch."""
    W, H = 1280, 720
    layout = bands(H, 6)
    scene = content.scene_pan(W, H, seed=2)
    frames = [scene.frame(i) for i in range(nframes)]
    per_band = [band_frames(f, W, H, layout) for f in frames]
    return {"sizes": [(W, bh) for _, bh in layout], "seq_frames": [[pb[b] for pb in per_band] for b in range(len(layout))],
            "qps": [24] * len(layout), "frames": frames,
            "desc": "1280x720 4:2:0 P-frames, QP 24, 8 tiles (6-MB-row bands)",
            "data": "synthetic (seeded pan + objects + grain)",
            "extra": {"frame": [W, H], "bands": len(layout), "qp": 24}}


def _job_batched(content, nframes):
    """16 independent streams in one batch: 8 x 640x368, 4 x 1280x720, 4 x
    1920x1088; moving textured rectangles over blurred noise with per-stream
    seeds; QP drawn from {20, 24, 28, 32} per stream (seeded)."""
    rng = np.random.default_rng(2024)
    sizes = [(640, 368)] * 8 + [(1280, 720)] * 4 + [(1920, 1088)] * 4
    qps = [int(q) for q in rng.choice([20, 24, 28, 32], size=len(sizes))]
    seq_frames = []
    for k, (w, h) in enumerate(sizes):
        sc = content.scene_objects(w, h, 300 + k, nobj=4 + w // 320, smin=24, smax=96, vmax=8, grain=1.0)
        seq_frames.append([sc.frame(i) for i in range(nframes)])
    return {"sizes": sizes, "seq_frames": seq_frames, "qps": qps, "frames": None,
            "desc": "16 streams (8 x 640x368, 4 x 1280x720, 4 x 1920x1088) in one batch, QP per stream "
                    "from {20,24,28,32}",
            "data": "synthetic (seeded moving rectangles over blurred noise, grain sigma 1)",
            "extra": {"qps": qps}}


def _job_stress(w, h):
    def make(content, nframes):
        """500 blobs of 4-12 px at velocities up to 48 px/frame, a new background
        every 40 frames, grain sigma 5; QP 24; 6-MB-row bands in one batch."""
        layout = bands(h, 6)
        sc = content.scene_blobs(w, h, 5)
        frames = [sc.frame(i) for i in range(nframes)]
        per_band = [band_frames(f, w, h, layout) for f in frames]
        return {"sizes": [(w, bh) for _, bh in layout],
                "seq_frames": [[pb[b] for pb in per_band] for b in range(len(layout))],
                "qps": [24] * len(layout), "frames": None,
                "desc": "%dx%d 4:2:0 P-frames, QP 24, %d tiles (6-MB-row bands), 500 blobs, scene cut every "
                        "40 frames" % (w, h, len(layout)),
                "data": "synthetic (seeded blobs, grain sigma 5)",
                "extra": {"frame": [w, h], "bands": len(layout), "qp": 24}}
    return make


WORKLOADS = {"720p": _job_720p, "batched": _job_batched, "stress1080": _job_stress(1920, 1088),
             "stress4k": _job_stress(3840, 2160)}


def run_steps(a, m, first, n, ncycle, qps, pipelined):
    """Frame steps first .. first + n - 1 of analyser a's batch in search mode m:
    step i codes source slot 1 + i % ncycle with one reference on step 0, two
    after.  Synchronous: each step queued, then its streams fetched.
    Pipelined: steps queued two deep from the resident sources
    (fdvgpu_submit_slots; a step's serialisation runs beside the next step's
    motion tables), each collected (its streams staged in host memory) once the
    next is queued.  Returns each step's stage times (ms)."""
    nb = a.nseq

    def args(i):
        return [1 + (i % ncycle)] * nb, qps, [1.0] * nb, [1 if i == 0 else 2] * nb, [0] * nb

    kt = []
    if not pipelined:
        for i in range(first, first + n):
            a.pframe(*args(i), m)
            a.fetch()
            kt.append(a.timings())
        return kt
    if n <= 0:
        return kt
    a.submit(0, *args(first), m)
    for k in range(n):
        if k + 1 < n:
            a.submit((k + 1) % 2, *args(first + k + 1), m)
        a.collect(k % 2)
        kt.append(a.slot_timings(k % 2))
    return kt


def bench_1080p(content, G, steps, warmup, pipelined=True):
    """Learned and exhaustive Mpix/s of 1920x1088 P-frames (the headline's pan
    content at 1080p), QP 24, as 12 bands of 6 MB rows in one batch."""
    w, h = 1920, 1088
    layout = bands(h, 6)
    sc = content.scene_pan(w, h, seed=2)
    nfr = steps + warmup + 1
    per = [band_frames(sc.frame(i), w, h, layout) for i in range(nfr)]
    sizes = [(w, bh) for _, bh in layout]
    nb = len(sizes)
    res = {}
    for name, m in (("learned", G.ME_LEARNED), ("exhaustive", G.ME_EXHAUSTIVE)):
        a = G.Analyser(sizes, nsrc=nfr)
        if m == G.ME_LEARNED:
            a.set_proposer()
        for b in range(nb):
            a.put_ref(b, 0, per[0][b])
            for i in range(1, nfr):
                a.put_source(b, i, per[i][b])
        a.autotune([1] * nb, [24] * nb, navail=[1] * nb, me_mode=m, reps=2)
        run_steps(a, m, 0, warmup, nfr - 1, [24] * nb, pipelined)
        t0 = time.perf_counter()
        run_steps(a, m, warmup, steps, nfr - 1, [24] * nb, pipelined)
        res[name] = steps * w * h / (time.perf_counter() - t0) / 1e6
        a.close()
    return {"frame": [w, h], "bands": nb, "qp": 24, "learned_mpix_s": round(res["learned"], 3),
            "exhaustive_mpix_s": round(res["exhaustive"], 3),
            "speedup_vs_exhaustive": round(res["learned"] / res["exhaustive"], 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="learned", choices=["diamond", "exhaustive", "learned", "parallel"])
    ap.add_argument("--workload", default="720p", choices=list(WORKLOADS),
                    help="720p: the headline workload (8 bands of one 1280x720 stream); batched: 16 streams "
                         "of mixed sizes; stress1080 / stress4k: 500 small fast blobs with scene cuts, "
                         "as 6-MB-row bands")
    ap.add_argument("--end-to-end", action="store_true",
                    help="also time the drop-in encoder end to end (source upload, device analysis, "
                         "stream fetch, host entropy coding) on the same frames")
    ap.add_argument("--stage", type=int, default=None, choices=[0, 1, 2, 3, 4],
                    help="structure of the block-parallel decision stage: 0 separate launches (costs, replays, "
                         "coding), 1 costs, replays, coding workers on one stream, 2 replays beside the costs and "
                         "workers beside the replays, 3 replays beside the costs, then the workers, 4 as 2 with "
                         "more workers joining once the replays are done (default: the autotuner's choice of 2, 3, "
                         "4)")
    ap.add_argument("--sync", action="store_true",
                    help="time frame steps one at a time (queue, then fetch the streams) instead of two deep "
                         "(each step's serialisation beside the next step's motion tables)")
    ap.add_argument("--conf", type=float, default=None, help="proposer confidence threshold (default 0.7)")
    ap.add_argument("--skip", type=float, default=None, help="proposer SKIP threshold (default: off; 0.7 prunes)")
    ap.add_argument("--baseline-steps", type=int, default=10,
                    help="steps of the exhaustive search timed for comparison (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--keyframe", action="store_true",
                    help="only time key frames (the device intra quadtree) at 720p, tiled and whole, beside the "
                         "reference's key-frame coder on the host")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (reads MIB, writes MIB): "
                         "the calibration for HBM byte counters")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(args.copy_probe << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_mib": args.copy_probe, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py: this encode-analysis path runs on one GPU; --gpus %d is not supported "
              "(independent streams or bands would shard across GPUs with no exchange)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from fdv_amd import content
    from fdv_amd import gpu as G

    if args.keyframe:
        bench_keyframe(args, content, G)
        return

    nsteps = args.warmup + args.steps
    nframes = min(nsteps + 1, 60)
    job = WORKLOADS[args.workload](content, nframes)
    W, H, QP, layout = 1280, 720, 24, bands(720, 6)
    frames = job.get("frames")
    mode = {"diamond": G.ME_DIAMOND, "exhaustive": G.ME_EXHAUSTIVE, "learned": G.ME_LEARNED,
            "parallel": G.ME_PARALLEL}[args.mode]

    sizes, per_seq, qps = job["sizes"], job["seq_frames"], job["qps"]
    nb = len(sizes)

    def open_analyser(m):
        a = G.Analyser(sizes, nsrc=nframes)
        if args.stage is not None:
            a.set_overlap(args.stage)
        if m == G.ME_LEARNED:
            a.set_proposer()
            if args.conf is not None or args.skip is not None:
                a.set_proposer_thresholds(0.7 if args.conf is None else args.conf,
                                          2.0 if args.skip is None else args.skip)
        for b in range(nb):
            a.put_ref(b, 0, per_seq[b][0])
            for i in range(1, nframes):
                a.put_source(b, i, per_seq[b][i])
        # launch shapes for this batch, timed on its first frame (the reference pool is restored)
        a.autotune([1] * nb, qps, navail=[1] * nb, me_mode=m, reps=2)
        return a

    an = open_analyser(mode)
    pipe = not args.sync
    run_steps(an, mode, 0, args.warmup, nframes - 1, qps, pipe)
    t0 = time.perf_counter()
    kt = run_steps(an, mode, args.warmup, args.steps, nframes - 1, qps, pipe)
    dt = time.perf_counter() - t0
    kt = np.array(kt)

    px_frame = sum(w * h for w, h in sizes)
    value = args.steps * px_frame / dt / 1e6
    k_ms = float(kt[:, 1].mean())                  # the decision stage (wavefront or block-parallel)
    bytes_per_px = 6.0                              # 1.5 B x (source + 2 references + recon)
    algo_bytes = bytes_per_px * px_frame
    achieved = algo_bytes / (k_ms * 1e-3) / 1e9
    traffic, traffic_src = profiled_traffic(args.mode) if args.workload == "720p" else (None, None)
    par = mode in (G.ME_LEARNED, G.ME_PARALLEL)
    # the dominant launch: the decision stage (k_pcosts, k_pxrep, k_pdecide: one graph launch)
    # in the block-parallel modes, the k_analyse wavefront otherwise; timed with
    # HIP events on the stream it runs on
    roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic, "traffic_source": traffic_src,
            "kernel": "decision stage (k_pcosts, k_pxrep, k_pdecide)" if par else "k_analyse",
            "kernel_ms": round(k_ms, 4), "bytes_per_launch": int(algo_bytes),
            "path_achieved": round(value * bytes_per_px / 1000.0, 3),
            "path_frac": round(value * bytes_per_px / 1000.0 / HBM_PEAK_GBS, 6)}
    stage_ms = {"motion_tables": round(float(kt[:, 0].mean()), 4), "decisions": round(k_ms, 4),
                "deblock": round(float(kt[:, 2].mean()), 4), "serialise": round(float(kt[:, 3].mean()), 4)}
    shapes = an.tune()
    structure = an.overlap() if par else None
    an.close()
    # the exhaustive search on the same frames, for the learned-vs-exhaustive comparison
    base = None
    if args.baseline_steps > 0 and mode != G.ME_EXHAUSTIVE:
        ab = open_analyser(G.ME_EXHAUSTIVE)
        run_steps(ab, G.ME_EXHAUSTIVE, 0, 2, nframes - 1, qps, pipe)
        tb = time.perf_counter()
        run_steps(ab, G.ME_EXHAUSTIVE, 2, args.baseline_steps, nframes - 1, qps, pipe)
        base = args.baseline_steps * px_frame / (time.perf_counter() - tb) / 1e6
        ab.close()
    # the parallel mode (exhaustive tables, the same block-parallel decisions,
    # no network) on the same frames: what the proposer adds or costs
    par_mpix = None
    if args.baseline_steps > 0 and mode == G.ME_LEARNED:
        ap_ = open_analyser(G.ME_PARALLEL)
        run_steps(ap_, G.ME_PARALLEL, 0, 2, nframes - 1, qps, pipe)
        tb = time.perf_counter()
        run_steps(ap_, G.ME_PARALLEL, 2, args.baseline_steps, nframes - 1, qps, pipe)
        par_mpix = args.baseline_steps * px_frame / (time.perf_counter() - tb) / 1e6
        ap_.close()
    # the 1080p half of the metric: 1920x1088 (the same pan content) as
    # 6-MB-row bands, learned and exhaustive on the same frames
    line1080 = None
    if args.baseline_steps > 0 and args.workload == "720p" and mode == G.ME_LEARNED:
        line1080 = bench_1080p(content, G, args.baseline_steps, min(args.warmup, 2), pipe)
    e2e = None
    if args.end_to_end:
        # the drop-in encoder (include/fdv_amd.h): every band's source uploaded,
        # analysed, its streams fetched and entropy-coded on host threads
        enc = G.Encoder(sizes, me_mode=mode)
        if mode == G.ME_LEARNED:
            enc.set_proposer()
        for b in range(nb):
            enc.set_ref(b, 0, per_seq[b][0])
        tcs = [G.TabCache() for _ in sizes]
        nbytes = 0

        def e2e_step(i):
            src = 1 + (i % (nframes - 1))
            pays, _ = enc.pframe_batch([per_seq[b][src] for b in range(nb)], [1 if i == 0 else 2] * nb, qps, tcs,
                                       recon=False)
            return sum(len(p) for p in pays)

        for i in range(args.warmup):
            e2e_step(i)
        te = time.perf_counter()
        for i in range(args.warmup, nsteps):
            nbytes += e2e_step(i)
        de = time.perf_counter() - te
        enc.close()
        e2e = {"value": round(args.steps * px_frame / de / 1e6, 3), "unit": "Mpix/s",
               "ms_per_step": round(de / args.steps * 1e3, 4),
               "payload_bytes_per_step": int(nbytes / args.steps),
               "includes": "source upload, device analysis, stream fetch, host entropy coding (one thread per "
                           "sequence, up to 16)"}
        # the same steps pipelined (fixed QP): frame k + 1 is analysed while the
        # host packs frame k (fdvamd_pframe_submit / fdvamd_pframe_collect)
        enc = G.Encoder(sizes, me_mode=mode)
        if mode == G.ME_LEARNED:
            enc.set_proposer()
        for b in range(nb):
            enc.set_ref(b, 0, per_seq[b][0])
        tcs = [G.TabCache() for _ in sizes]

        def submit(i):
            src = 1 + (i % (nframes - 1))
            enc.pframe_submit([per_seq[b][src] for b in range(nb)], [1 if i == 0 else 2] * nb, qps)

        def collect():
            return sum(len(p) for p in enc.pframe_collect(tcs))

        submit(0)
        for i in range(1, args.warmup):
            submit(i)
            collect()
        te = time.perf_counter()
        pbytes = 0
        for i in range(args.warmup, nsteps):
            submit(i)
            pbytes += collect()
        pbytes += collect()
        dp = time.perf_counter() - te
        enc.close()
        e2e["pipelined"] = {"value": round(args.steps * px_frame / dp / 1e6, 3), "unit": "Mpix/s",
                            "ms_per_step": round(dp / args.steps * 1e3, 4),
                            "payload_bytes_per_step": int(pbytes / args.steps)}
    # BD-rate of this search mode against the exhaustive one, from the committed
    # rate-distortion study over the reference's scene library (bdrate.py --library)
    bd = None
    bd_path = os.path.join(ROOT, "fdv_amd", "learned", "bdrate_library_720p.json")
    if mode in (G.ME_LEARNED, G.ME_PARALLEL) and os.path.exists(bd_path):
        with open(bd_path) as f:
            rec = json.load(f)[args.mode]
        bd = {"median_percent": round(rec["summary"]["median"], 2), "mean_percent": round(rec["summary"]["mean"], 2),
              "scenes": rec["summary"]["scenes"], "worst": rec["summary"]["worst"],
              "qps": rec["qps"], "source": "fdv_amd/learned/bdrate_library_720p.json"}
        # the same study at the proposer configuration's QPs, and on seeded content
        for key, name in (("library_qp20_32", "bdrate_library_720p_qp20-32.json"),
                          ("seeded", "bdrate_seeded_720p.json")):
            pth = os.path.join(ROOT, "fdv_amd", "learned", name)
            if os.path.exists(pth):
                with open(pth) as f:
                    r2 = json.load(f).get(args.mode)
                if r2 and r2.get("summary"):
                    bd[key] = {"median_percent": round(r2["summary"]["median"], 2),
                               "mean_percent": round(r2["summary"]["mean"], 2), "qps": r2["qps"],
                               "scenes": r2["summary"]["scenes"], "worst": r2["summary"]["worst"],
                               "source": "fdv_amd/learned/" + name}
    # the CPU leg runs on the headline workload only
    cpu = None if (args.no_cpu or args.workload != "720p") else cpu_baseline(frames, W, layout, QP, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": job["data"],
        "config": {"workload": job["desc"] + ", %s motion search%s" % (
                       args.mode, ", block-parallel decisions" if mode in (G.ME_LEARNED, G.ME_PARALLEL) else ""),
                   "sequences": nb, "coded_px_per_step": px_frame, "search": args.mode,
                   "steps_in_flight": 1 if args.sync else 2, **job["extra"]},
        "stage_ms": stage_ms,
        "launch_shapes": {"serialiser_waves": shapes[0], "deblock_threads": shapes[1], "proposer_wg_per_cu": shapes[2],
                          "decision_workers_per_cu": shapes[3], "decision_structure": structure},
        "exhaustive_mpix_s": None if base is None else round(base, 3),
        "speedup_vs_exhaustive": None if base is None else round(value / base, 3),
        "parallel_mpix_s": None if par_mpix is None else round(par_mpix, 3),
        "speedup_vs_parallel": None if par_mpix is None else round(value / par_mpix, 3),
        "line_1080p": line1080,
        "bd_rate_vs_exhaustive": bd,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if e2e is not None:
        out["end_to_end"] = e2e
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
