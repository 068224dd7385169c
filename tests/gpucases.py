"""Shared drivers for the GPU parity tests: run a batch of sequences through the
HIP path and through the oracle, frame by frame, and compare everything."""
from __future__ import annotations

import sys

import numpy as np

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mbinfo_to_np(info):
    out = np.zeros(len(info), dtype=G.MBINFO_DTYPE)
    for i, m in enumerate(info):
        out[i]["mode"] = m.mode
        out[i]["ref"] = m.ref
        out[i]["merge_idx"] = m.merge_idx
        out[i]["sub"] = m.sub
        out[i]["mv"] = np.array([[m.mv[k][0], m.mv[k][1]] for k in range(4)])
    return out


def np_to_mbinfo(arr):
    info = (L.MbInfo * len(arr))()
    for i, m in enumerate(arr):
        info[i].mode = int(m["mode"]); info[i].ref = int(m["ref"])
        info[i].merge_idx = int(m["merge_idx"]); info[i].sub = int(m["sub"])
        for k in range(4):
            info[i].mv[k][0] = int(m["mv"][k][0]); info[i].mv[k][1] = int(m["mv"][k][1])
    return info


def same_decisions(a, b, nmb):
    """Compare decisions on the fields each mode defines."""
    for i in range(nmb):
        x, y = a[i], b[i]
        if x["mode"] != y["mode"]:
            return i
        m = int(x["mode"])
        if m == 0 and (x["merge_idx"] != y["merge_idx"] or x["ref"] != y["ref"] or
                       not np.array_equal(x["mv"][0], y["mv"][0])):
            return i
        if m == 1 and (x["ref"] != y["ref"] or not np.array_equal(x["mv"][0], y["mv"][0])):
            return i
        if m == 2 and x["sub"] != y["sub"]:
            return i
        if m == 3 and not np.array_equal(x["mv"], y["mv"]):
            return i
    return -1


def run_parity(scenes, nframes, qps, me_mode=G.ME_DIAMOND, lscales=None, force_skip_at=None,
               given=False, label=""):
    """scenes: list of content.Scene.  Frame 0 is installed as the reference (an
    ideal key frame); frames 1.. are P-frames chained through the GPU's own
    reconstruction.  Returns the number of mismatching (sequence, frame) pairs."""
    n = len(scenes)
    sizes = [(s.w, s.h) for s in scenes]
    lscales = lscales or [1.0] * n
    an = G.Analyser(sizes, nsrc=1)
    frames = [[s.frame(i) for i in range(nframes)] for s in scenes]
    near = [frames[k][0] for k in range(n)]
    far = [None] * n
    nref = [1] * n
    for k in range(n):
        an.put_ref(k, 0, near[k])
    bad = 0
    for f in range(1, nframes):
        navail = [2 if nref[k] >= 2 else 1 for k in range(n)]
        fs = [1 if (force_skip_at is not None and f == force_skip_at) else 0 for _ in range(n)]
        orc = []
        for k in range(n):
            w, h = sizes[k]
            if me_mode in (G.ME_PARALLEL, G.ME_LEARNED) and not fs[k] and not given:
                st, rec, info = L.oracle_parallel(frames[k][f], near[k], far[k], w, h, navail[k], qps[k],
                                                        lscales[k])
            else:
                st, rec, info = L.oracle_analyse(frames[k][f], near[k], far[k], w, h, navail[k], qps[k],
                                                 lscales[k], fs[k], 0 if me_mode == G.ME_DIAMOND else 1)
            orc.append((st, rec, info))
            an.put_source(k, 0, frames[k][f])
            if given:
                an.set_given(k, mbinfo_to_np(info))
        say(f"  [{label}] frame {f}: GPU pframe over {n} sequence(s) sizes={sizes} qp={qps} me={me_mode}")
        an.pframe([0] * n, qps, lscales, navail, fs, me_mode)
        an.fetch()
        for k in range(n):
            w, h = sizes[k]
            st, rec, info = orc[k]
            gst = an.streams(k)
            grec = an.recon(k)
            gdec = an.decisions(k)
            odec = mbinfo_to_np(info)
            nmb = (w // 16) * (h // 16)
            di = same_decisions(gdec, odec, nmb)
            sok = all(gst[j] == st[j] for j in range(16))
            rok = np.array_equal(grec, rec)
            if di >= 0 or not sok or not rok:
                bad += 1
                bads = [j for j in range(16) if gst[j] != st[j]]
                say(f"  MISMATCH seq {k} frame {f}: first decision diff MB {di} "
                    f"(gpu {gdec[di] if di >= 0 else None} oracle {odec[di] if di >= 0 else None}); "
                    f"streams differ {bads}; recon equal {rok}")
            # chain through the oracle's reconstruction (identical when parity holds)
            far[k] = near[k]
            near[k] = rec
            nref[k] = min(nref[k] + 1, 2)
        if bad:
            # resync the GPU pool with the oracle's references before continuing
            for k in range(n):
                an.put_ref(k, 0, near[k])
                if far[k] is not None:
                    an.put_ref(k, 1, far[k])
    an.close()
    return bad
