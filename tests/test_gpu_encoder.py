"""The drop-in encoder (include/fdv_amd.h) on the MI355X path against the
oracle's pframe_encode: identical payload bytes, reconstruction and carried
entropy state over chains of P-frames, one sequence at a time and batched; and
the 1280x720 eight-band workload (the reference's tiled encoder with 6-MB-row
bands, fdv.h:7161) through the analysis path at full size."""
import numpy as np
import pytest

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import mbinfo_to_np, run_parity, say

pytestmark = pytest.mark.gpu


def _oracle_chain(frames, w, h, qp, lscale=1.0, fs_at=None):
    tc = L.TabCache()
    near, far = frames[0], None
    out = []
    for f in range(1, len(frames)):
        navail = 1 if f == 1 else 2
        pay, rec, _ = L.oracle_pframe(frames[f], near, far, w, h, navail, qp, tc, lscale,
                                      1 if fs_at == f else 0)
        out.append((pay.tobytes(), rec, tc.buf.copy()))
        far, near = near, rec
    return out


def test_pframe_encode_chain_matches_oracle():
    sc = content.scene_pan(96, 64, 21, nobj=3, smin=8, smax=32)
    frames = [sc.frame(i) for i in range(6)]
    want = _oracle_chain(frames, sc.w, sc.h, 26)
    enc = G.Encoder([(sc.w, sc.h)])
    enc.set_ref(0, 0, frames[0])
    tc = G.TabCache()
    for f in range(1, 6):
        say(f"  dropin pframe_encode 96x64 frame {f}")
        pay, rec = enc.pframe_encode(frames[f], 1 if f == 1 else 2, 26, tc)
        wp, wr, wt = want[f - 1]
        assert pay == wp, f"frame {f}: payload {len(pay)} B vs oracle {len(wp)} B"
        assert np.array_equal(rec, wr), f"frame {f}: recon"
        assert np.array_equal(tc.buf, wt), f"frame {f}: entropy state"
    enc.close()


def test_pframe_batch_mixed_sizes_matches_oracle():
    scenes = [content.scene_objects(64, 48, 22, nobj=2, smin=8, smax=24),
              content.scene_pan(128, 64, 23, nobj=3, smin=8, smax=40),
              content.scene_objects(320, 180, 24)]
    qps = [20, 32, 28]
    ls = [1.0, 0.8, 1.0]
    nfr = 4
    frames = [[s.frame(i) for i in range(nfr)] for s in scenes]
    want = [_oracle_chain(frames[k], s.w, s.h, qps[k], ls[k], fs_at=2 if k == 0 else None)
            for k, s in enumerate(scenes)]
    enc = G.Encoder([(s.w, s.h) for s in scenes])
    for k in range(3):
        enc.set_ref(k, 0, frames[k][0])
    tcs = [G.TabCache() for _ in scenes]
    for f in range(1, nfr):
        say(f"  dropin batch of 3 sequences frame {f}")
        pays, recs = enc.pframe_batch([frames[k][f] for k in range(3)], [1 if f == 1 else 2] * 3, qps, tcs,
                                      ls, [1 if (k == 0 and f == 2) else 0 for k in range(3)])
        for k in range(3):
            wp, wr, wt = want[k][f - 1]
            assert pays[k] == wp, f"seq {k} frame {f}: payload"
            assert np.array_equal(recs[k], wr), f"seq {k} frame {f}: recon"
            assert np.array_equal(tcs[k].buf, wt), f"seq {k} frame {f}: entropy state"
    enc.close()


class _Band:
    """One 6-MB-row band of a 1280x720 scene as a sequence of its own."""

    def __init__(self, scene, y0, bh):
        self.scene, self.y0, self.w, self.h = scene, y0, scene.w, bh
        self._cache = {}

    def frame(self, i):
        if i not in self._cache:
            y, u, v = self.scene.planes(i)
            y0, bh = self.y0, self.h
            self._cache[i] = np.concatenate([y[y0:y0 + bh].ravel(), u[y0 // 2:(y0 + bh) // 2].ravel(),
                                             v[y0 // 2:(y0 + bh) // 2].ravel()])
        return self._cache[i]


def test_720p_eight_bands_full_size():
    sc = content.scene_pan(1280, 720, 2)
    bands = [_Band(sc, y0, min(96, 720 - y0)) for y0 in range(0, 720, 96)]
    assert [b.h for b in bands] == [96] * 7 + [48]
    assert run_parity(bands, 3, [24] * 8, label="720p-bands") == 0


def test_720p_eight_bands_full_size_parallel_mode():
    """The block-parallel decisions at their real shape: every band's table,
    decisions, streams and recon equal to the oracle's MB loop on the table
    over a chain of P-frames."""
    sc = content.scene_pan(1280, 720, 4)
    bands = [_Band(sc, y0, min(96, 720 - y0)) for y0 in range(0, 720, 96)]
    assert run_parity(bands, 4, [24] * 8, me_mode=G.ME_PARALLEL, label="720p-parallel") == 0


def test_stress_blobs_parallel_mode():
    """Fast small blobs with a scene cut and heavy grain (the stress content at
    640x368): many intra MBs, so long chains through the intra pass's handoff
    words, and every decision path of the replay; equal to the oracle."""
    sc = content.scene_blobs(640, 368, 7, nblob=120, cut=3)
    bands = [_Band(sc, y0, min(96, 368 - y0)) for y0 in range(0, 368, 96)]
    assert run_parity(bands, 5, [20, 28, 34, 40], me_mode=G.ME_PARALLEL, label="stress-parallel") == 0


def test_720p_eight_bands_given_reference_decisions():
    sc = content.scene_pan(1280, 720, 3)
    bands = [_Band(sc, y0, min(96, 720 - y0)) for y0 in range(0, 720, 96)]
    assert run_parity(bands, 3, [24] * 8, given=True, label="720p-given") == 0


def test_requantising_retry_after_rewind():
    """The streaming encoder codes a frame twice when the first attempt misses
    its budget (fdv.h:8303-8334): the second attempt starts from the same
    references and the entropy state from before the first."""
    sc = content.scene_objects(96, 64, 31, nobj=3, smin=8, smax=32)
    frames = [sc.frame(i) for i in range(4)]
    enc = G.Encoder([(sc.w, sc.h)])
    enc.set_ref_planes(0, 0, frames[0])
    tc, tco = G.TabCache(), L.TabCache()
    near, far = frames[0], None
    for f in range(1, 4):
        navail = 1 if f == 1 else 2
        before = tc.buf.copy()
        say(f"  retry: frame {f} at qp 20, rewind, again at qp 34")
        enc.pframe_encode(frames[f], navail, 20, tc)
        enc.rewind()
        tc.buf[:] = before
        pay, rec = enc.pframe_encode(frames[f], navail, 34, tc)
        wp, wr, _ = L.oracle_pframe(frames[f], near, far, sc.w, sc.h, navail, 34, tco)
        assert pay == wp.tobytes(), f"frame {f}: payload"
        assert np.array_equal(rec, wr), f"frame {f}: recon"
        far, near = near, wr
    with pytest.raises(G.GpuError):
        enc.rewind()
        enc.rewind()
    enc.close()


@pytest.mark.parametrize("mode", [G.ME_DIAMOND, G.ME_PARALLEL, G.ME_LEARNED])
def test_pipelined_submit_collect_matches_sequential(mode):
    """Frames submitted two at a time and collected while the next one is
    analysed give the payloads of the synchronous batch call, frame by frame;
    misuse (a third frame in flight, a collect with none, a synchronous call
    while frames are in flight) fails without disturbing the pipeline."""
    scenes = [content.scene_objects(64, 48, 31, nobj=2, smin=8, smax=24),
              content.scene_pan(128, 64, 32, nobj=3, smin=8, smax=40)]
    qps = [24, 30]
    nfr = 6
    frames = [[s.frame(i) for i in range(nfr + 3)] for s in scenes]
    sizes = [(s.w, s.h) for s in scenes]

    def fresh():
        e = G.Encoder(sizes, me_mode=mode)
        if mode == G.ME_LEARNED:
            e.set_proposer()
        for k in range(2):
            e.set_ref(k, 0, frames[k][0])
        return e, [G.TabCache() for _ in scenes]

    def fr(f):
        return [frames[k][f] for k in range(2)]

    def nav(f):
        return [1 if f == 1 else 2] * 2

    say(f"  pipelined encoder, mode {mode}: synchronous chain")
    a, ta = fresh()
    want = [a.pframe_batch(fr(f), nav(f), qps, ta, recon=False)[0] for f in range(1, nfr)]
    a.close()
    say(f"  pipelined encoder, mode {mode}: submit / collect chain")
    b, tb = fresh()
    got = []
    b.pframe_submit(fr(1), nav(1), qps)
    for f in range(2, nfr):
        b.pframe_submit(fr(f), nav(f), qps)
        got.append(b.pframe_collect(tb))
    got.append(b.pframe_collect(tb))
    for f in range(len(want)):
        assert got[f] == want[f], f"frame {f + 1}: pipelined payloads differ"
    with pytest.raises(G.GpuError):
        b.pframe_collect(tb)
    b.pframe_submit(fr(nfr), nav(nfr), qps)
    b.pframe_submit(fr(nfr + 1), nav(nfr + 1), qps)
    with pytest.raises(G.GpuError):
        b.pframe_submit(fr(nfr + 2), nav(nfr + 2), qps)
    with pytest.raises(G.GpuError):
        b.pframe_batch(fr(nfr + 2), nav(nfr + 2), qps, tb, recon=False)
    assert all(len(p) > 0 for p in b.pframe_collect(tb))
    assert all(len(p) > 0 for p in b.pframe_collect(tb))
    b.close()


def test_pframe_encode_rejects_a_batch_encoder():
    """fdvamd_pframe_encode codes sequence 0 of a one-sequence encoder; with
    more sequences it fails with an error instead of reading past its
    one-element argument arrays."""
    enc = G.Encoder([(64, 48), (64, 48)])
    enc.set_ref(0, 0, np.zeros(64 * 48 * 3 // 2, np.uint8))
    with pytest.raises(G.GpuError, match="one-sequence"):
        enc.pframe_encode(np.zeros(64 * 48 * 3 // 2, np.uint8), 1, 24)
    enc.close()


def test_wavefront_tags_survive_many_block_parallel_frames():
    """The exact wavefront's handoff words are tagged per frame of that kernel
    only: a given-decisions frame 255 block-parallel frames after the previous
    one still waits for this frame's neighbours (a stale tag would let it read
    the old frame's merge candidates and reconstruction)."""
    sc = content.scene_objects(64, 64, 77, nobj=4, smin=8, smax=24, vmax=6, grain=1.0)
    w, h = sc.w, sc.h
    nfr = 258
    an = G.Analyser([(w, h)], nsrc=1)
    an.set_proposer()
    an.put_ref(0, 0, sc.frame(0))
    recs = [sc.frame(0)]
    say(f"  tag test: {nfr - 1} frames, given decisions at 1 and {nfr - 1}")
    for f in range(1, nfr):
        cur = sc.frame(f % 8 + 1)
        navail = 1 if f == 1 else 2
        given = f in (1, nfr - 1)
        if given:
            _, _, info = L.oracle_analyse(cur, recs[-1], recs[-2] if navail == 2 else None, w, h, navail, 24)
            gv = mbinfo_to_np(info)
            assert (gv["mode"] == 0).any() or (gv["mode"] == 2).any()
            an.set_given(0, gv)
        an.put_source(0, 0, cur)
        an.pframe([0], [24], navail=[navail], me_mode=G.ME_LEARNED)
        an.fetch()
        rec = an.recon(0)
        if given:
            st, orec, _ = L.oracle_analyse(cur, recs[-1], recs[-2] if navail == 2 else None, w, h, navail, 24,
                                           given=info)
            assert an.streams(0) == st, f"frame {f}: streams"
            assert np.array_equal(rec, orec), f"frame {f}: reconstruction"
        recs = recs[-1:] + [rec]
    an.close()


@pytest.mark.parametrize("mode", [G.ME_EXHAUSTIVE, G.ME_PARALLEL, G.ME_LEARNED])
def test_retry_reuses_the_motion_tables(mode):
    """A requantising retry after fdvgpu_rewind_keep_motion keeps the first
    attempt's motion tables, and its result is the oracle's at the retry QP
    from those tables (the exhaustive tables do not depend on the QP, so for
    them this is a fresh step's result)."""
    scenes = [content.scene_pan(320, 192, 35, nobj=4, smin=16, smax=48),
              content.scene_objects(128, 48, 36, nobj=3, smin=8, smax=32, vmax=8)]
    sizes = [(s.w, s.h) for s in scenes]
    frames = [[s.frame(i) for i in range(4)] for s in scenes]
    an = G.Analyser(sizes, nsrc=1)
    if mode == G.ME_LEARNED:
        an.set_proposer()
    near = [fr[0] for fr in frames]
    far = [None, None]
    for k in range(2):
        an.put_ref(k, 0, near[k])
    for f in range(1, 4):
        navail = [1 if f == 1 else 2] * 2
        for k in range(2):
            an.put_source(k, 0, frames[k][f])
        say(f"  retry mode {mode} frame {f}: qp 20, rewind keeping motion, qp 34")
        an.pframe([0, 0], [20, 20], navail=navail, me_mode=mode)
        an.fetch()
        tables = [an.motion(k) for k in range(2)]
        props = [an.proposals(k) for k in range(2)] if mode == G.ME_LEARNED else None
        an.rewind(keep_motion=True)
        an.pframe([0, 0], [34, 34], navail=navail, me_mode=mode)
        an.fetch()
        for k in range(2):
            w, h = sizes[k]
            assert an.motion(k).tobytes() == tables[k].tobytes()
            if mode == G.ME_LEARNED:
                mt = L.oracle_metab_candidates(frames[k][f], near[k], far[k], w, h, navail[k], props[k])
            else:
                mt = L.oracle_metab_exhaustive(frames[k][f], near[k], far[k], w, h, navail[k])
            if mode == G.ME_EXHAUSTIVE:
                st, rec, _ = L.oracle_analyse(frames[k][f], near[k], far[k], w, h, navail[k], 34, me_mode=1)
            else:
                st, rec, _ = L.oracle_parallel(frames[k][f], near[k], far[k], w, h, navail[k], 34, metab=mt)
            assert an.streams(k) == st, f"seq {k} frame {f}: streams"
            grec = an.recon(k)
            assert np.array_equal(grec, rec), f"seq {k} frame {f}: recon"
            far[k], near[k] = near[k], grec
    an.close()


def test_retry_skips_the_search_time():
    """At 1280x720 the exhaustive search is most of the tables stage: a retry
    that keeps the motion tables spends a fraction of a first attempt's time
    there (the median of several retries against several first attempts)."""
    sc = content.scene_pan(1280, 720, 37)
    frames = [sc.frame(i) for i in range(3)]
    an = G.Analyser([(1280, 720)], nsrc=2)
    an.put_ref(0, 0, frames[0])
    an.put_source(0, 0, frames[1])
    first, retry = [], []
    for rep in range(6):
        an.pframe([0], [20], navail=[1], me_mode=G.ME_EXHAUSTIVE)
        an.fetch()
        first.append(an.timings()[0])
        an.rewind(keep_motion=True)
        an.pframe([0], [34], navail=[1], me_mode=G.ME_EXHAUSTIVE)
        an.fetch()
        retry.append(an.timings()[0])
        an.rewind()
    f, r = sorted(first)[3], sorted(retry)[3]
    say(f"  tables stage: first attempt {f:.4f} ms, retry {r:.4f} ms")
    assert r < 0.5 * f
    an.close()
