"""Frame steps replayed from captured hipGraphs, and autotuned launch shapes,
give exactly the results of eager launches: every replay (with eager work and
synchronisation between replays) against a context that never captures."""
import numpy as np
import pytest

from fdv_amd import content
from fdv_amd import gpu as G
import orclib as L
from gpucases import say

pytestmark = pytest.mark.gpu


def _scenes():
    return [content.scene_objects(96, 64, 61, nobj=3, smin=8, smax=32, vmax=6),
            content.scene_pan(128, 48, 62, nobj=2, smin=8, smax=24, pan=(3, -1), grain=2.0)]


@pytest.mark.parametrize("mode", [G.ME_DIAMOND, G.ME_PARALLEL, G.ME_LEARNED])
def test_graph_replays_match_eager(mode):
    scenes = _scenes()
    sizes = [(s.w, s.h) for s in scenes]
    nfr = 6
    frames = [[s.frame(i) for i in range(nfr)] for s in scenes]
    ctx = []
    for graphs in (False, True):
        an = G.Analyser(sizes, nsrc=nfr)
        an.set_graphs(graphs)
        if mode == G.ME_LEARNED:
            an.set_proposer()
        for k in range(2):
            an.put_ref(k, 0, frames[k][0])
            for f in range(1, nfr):
                an.put_source(k, f, frames[k][f])
        ctx.append(an)
    for f in range(1, nfr):
        say(f"  graph replay mode {mode} frame {f}")
        outs = []
        for an in ctx:
            an.pframe([f, f], [24, 30], navail=[1 if f == 1 else 2] * 2, me_mode=mode)
            an.fetch()
            outs.append([(an.streams(k), an.recon(k)) for k in range(2)])
            np.sort(np.random.default_rng(f).random(10000))      # eager host work between replays
        for k in range(2):
            assert outs[0][k][0] == outs[1][k][0], f"frame {f} seq {k}: streams"
            assert np.array_equal(outs[0][k][1], outs[1][k][1]), f"frame {f} seq {k}: recon"
    for an in ctx:
        an.close()


@pytest.mark.parametrize("mode", [G.ME_DIAMOND, G.ME_PARALLEL, G.ME_LEARNED])
def test_pipelined_steps_match_synchronous(mode):
    """Frame steps queued two deep from the source slots (fdvgpu_submit_slots:
    the serialisation beside the next step's tables, on a stream of its own)
    give every frame's streams and the final reconstruction of steps run one at
    a time; a plain fetch after them returns the last one's streams, and a
    synchronous step after them is right too."""
    scenes = _scenes()
    sizes = [(s.w, s.h) for s in scenes]
    nfr = 7
    frames = [[s.frame(i) for i in range(nfr)] for s in scenes]
    ctx = []
    for _ in range(2):
        an = G.Analyser(sizes, nsrc=nfr)
        if mode == G.ME_LEARNED:
            an.set_proposer()
        for k in range(2):
            an.put_ref(k, 0, frames[k][0])
            for f in range(1, nfr):
                an.put_source(k, f, frames[k][f])
        ctx.append(an)
    sync, pipe = ctx
    want = []
    for f in range(1, nfr):
        sync.pframe([f, f], [24, 30], navail=[1 if f == 1 else 2] * 2, me_mode=mode)
        sync.fetch()
        want.append([sync.streams(k) for k in range(2)])
    last = nfr - 2                      # frames 1 .. nfr - 2 pipelined, the last one synchronous
    say(f"  pipelined mode {mode}: frames 1..{last}")
    pipe.submit(0, [1, 1], [24, 30], navail=[1, 1], me_mode=mode)
    for f in range(1, last + 1):
        if f < last:
            pipe.submit(f % 2, [f + 1] * 2, [24, 30], navail=[2, 2], me_mode=mode)
        pipe.collect((f - 1) % 2)
        for k in range(2):
            assert pipe.collected_streams((f - 1) % 2, k) == want[f - 1][k], f"frame {f} seq {k}: streams"
        assert all(t >= 0.0 for t in pipe.slot_timings((f - 1) % 2))
        np.sort(np.random.default_rng(f).random(10000))          # host work while the next step runs
    # a plain fetch after the pipelined steps returns the last one's streams
    pipe.fetch()
    for k in range(2):
        assert pipe.streams(k) == want[last - 1][k], f"fetch after the pipelined steps, seq {k}"
    pipe.pframe([nfr - 1] * 2, [24, 30], navail=[2, 2], me_mode=mode)
    pipe.fetch()
    for k in range(2):
        assert pipe.streams(k) == want[nfr - 2][k], f"synchronous frame after the pipelined ones, seq {k}"
        assert np.array_equal(pipe.recon(k), sync.recon(k)), f"seq {k}: recon"
    for an in ctx:
        an.close()


def test_graph_cache_keeps_stages_apart():
    """Each (stage, mode, decision structure) has its own captured graph: after
    learned frames whose decision stage ran from a graph in structure 3, a held
    frame in the parallel mode (which takes the MB wavefront) gives the result
    of a context that never captures (the two once shared a cache slot)."""
    scenes = _scenes()
    sizes = [(s.w, s.h) for s in scenes]
    nfr = 6
    frames = [[s.frame(i) for i in range(nfr)] for s in scenes]
    outs = []
    for graphs in (False, True):
        an = G.Analyser(sizes, nsrc=nfr)
        an.set_graphs(graphs)
        an.set_overlap(3)
        an.set_proposer()
        for k in range(2):
            an.put_ref(k, 0, frames[k][0])
            for f in range(1, nfr):
                an.put_source(k, f, frames[k][f])
        for f in range(1, 4):
            an.pframe([f, f], [24, 30], navail=[1 if f == 1 else 2] * 2, me_mode=G.ME_LEARNED)
            an.fetch()
        say(f"  graph cache: held parallel-mode frame, graphs {graphs}")
        an.pframe([4, 4], [24, 30], navail=[2, 2], force_skip=[1, 1], me_mode=G.ME_PARALLEL)
        an.fetch()
        outs.append([(an.streams(k), an.recon(k)) for k in range(2)])
        an.close()
    for k in range(2):
        assert outs[0][k][0] == outs[1][k][0], f"seq {k}: streams"
        assert np.array_equal(outs[0][k][1], outs[1][k][1]), f"seq {k}: recon"


def test_autotune_changes_no_result():
    scenes = _scenes()
    sizes = [(s.w, s.h) for s in scenes]
    frames = [[s.frame(i) for i in range(4)] for s in scenes]
    a, b = G.Analyser(sizes, nsrc=4), G.Analyser(sizes, nsrc=4)
    for an in (a, b):
        an.set_proposer()
        for k in range(2):
            an.put_ref(k, 0, frames[k][0])
            for f in range(1, 4):
                an.put_source(k, f, frames[k][f])
    say("  autotune")
    shapes = b.autotune([1, 1], [24, 30], navail=[1, 1], me_mode=G.ME_LEARNED, reps=1)
    assert shapes[0] in (2, 4, 8) and shapes[1] in (64, 128, 256) and shapes[2] in (1, 2) and shapes[3] in (1, 2)
    for f in range(1, 4):
        for an in (a, b):
            an.pframe([f, f], [24, 30], navail=[1 if f == 1 else 2] * 2, me_mode=G.ME_LEARNED)
            an.fetch()
        for k in range(2):
            assert a.streams(k) == b.streams(k) and np.array_equal(a.recon(k), b.recon(k)), f"frame {f} seq {k}"
    a.close()
    b.close()


@pytest.mark.parametrize("mode", [G.ME_PARALLEL, G.ME_LEARNED])
def test_fused_decision_stage_matches_separate_launches(mode):
    """The decision stage's five structures give the same results, eagerly and
    from captured graphs: separate launches (costs, replays, coding kernels);
    the costs, the replays, then the coding workers on one stream; the replays
    beside the costs on a second stream and the coding workers beside the
    replays; the replays beside the costs, then the coding workers; the
    workers beside the replays joined by more once the replays are done
    (tagged words throughout).  Each replay is one MB row; the third
    sequence has 11 rows.  In the parallel mode the result is also the oracle's
    MB loop on the exhaustive table, frame by frame."""
    scenes = _scenes() + [content.scene_pan(80, 176, 63, nobj=2, smin=8, smax=24, pan=(2, 1), grain=2.0)]
    sizes = [(s.w, s.h) for s in scenes]
    nfr = 5
    frames = [[s.frame(i) for i in range(nfr)] for s in scenes]
    outs = []
    for overlap in (0, 1, 2, 3, 4):
        an = G.Analyser(sizes, nsrc=nfr)
        an.set_overlap(overlap)
        if mode == G.ME_LEARNED:
            an.set_proposer()
        ns = len(scenes)
        for k in range(ns):
            an.put_ref(k, 0, frames[k][0])
            for f in range(1, nfr):
                an.put_source(k, f, frames[k][f])
        res = []
        for f in range(1, nfr):
            say(f"  fused {overlap} mode {mode} frame {f}")
            an.pframe([f] * ns, [24, 30, 27], navail=[1 if f == 1 else 2] * ns, me_mode=mode)
            an.fetch()
            res.append([(an.streams(k), an.recon(k).copy()) for k in range(ns)])
        an.close()
        outs.append(res)
    if mode == G.ME_PARALLEL:
        for k, sc in enumerate(scenes):
            near, far = frames[k][0], None
            for f in range(1, nfr):
                st, rec, _ = L.oracle_parallel(frames[k][f], near, far, sc.w, sc.h, 1 if f == 1 else 2,
                                               [24, 30, 27][k])
                assert outs[0][f - 1][k][0] == st, f"frame {f} seq {k}: streams differ from the oracle"
                assert np.array_equal(outs[0][f - 1][k][1], rec), f"frame {f} seq {k}: recon differs from the oracle"
                far, near = near, rec
    for v in (1, 2, 3, 4):
        for f in range(nfr - 1):
            for k in range(len(scenes)):
                assert outs[0][f][k][0] == outs[v][f][k][0], f"structure {v} frame {f + 1} seq {k}: streams"
                assert np.array_equal(outs[0][f][k][1], outs[v][f][k][1]), f"structure {v} frame {f + 1} seq {k}: recon"
