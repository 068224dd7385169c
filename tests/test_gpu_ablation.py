"""What the trained proposer gains over the search without it.

Both searches get the same budget and the same thresholds: the MBs whose
network confidence reaches 0.7 take a candidate search over the 3x3 whole-pel
neighbourhoods of K vectors (and the zero vector); the rest take the
exhaustive search.  With the network, the K are its top-K vectors; without it,
the K lowest positions of the proposer's own correlation map (ties to the
smaller |dx| + |dy|, then the lower position, as the map's minima are chosen).
The oracle's candidate search (orc_metab_candidates) evaluates both, and each
is scored against the exhaustive table over the candidate-searched MBs: how
many find the exhaustive 16x16 vector, and the SAD given away (the sum of
their 16x16 SAD above the exhaustive search's).

Measured on four seeded 640x368 scenes: the network gives away less SAD than
the map's minima with one candidate and with eight; with two and four the map's
minima give away less, and they find the exhaustive vector about as often
(within 1%) at every K.  The test holds the network to its measured gains.
"""
import numpy as np
import pytest
import torch

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import say

pytestmark = pytest.mark.gpu


def _maps(an, w, h, qp):
    """Per MB, the correlation map's positions in the order the proposer's
    minima use: relative cost, then |dx| + |dy|, then position."""
    lib = G.lib()
    nmb = (w // 16) * (h // 16)
    cy, _, ry, rs = an.planes(0, 0)
    pos = torch.tensor([[(i % (w // 16)) * 16, (i // (w // 16)) * 16] for i in range(nmb)], dtype=torch.int32,
                       device="cuda")
    # every argument array is bound to a name: a temporary's memory could be
    # recycled for the next one before the kernel reads it
    curs = torch.full((nmb,), cy, dtype=torch.int64, device="cuda")
    refs = torch.full((nmb,), ry, dtype=torch.int64, device="cuda")
    cw = torch.full((nmb,), w, dtype=torch.int32, device="cuda")
    rss = torch.full((nmb,), rs, dtype=torch.int32, device="cuda")
    qps = torch.full((nmb,), qp, dtype=torch.int32, device="cuda")
    feats = torch.zeros((nmb, 289 * 2), dtype=torch.float32, device="cuda")
    assert lib.fdvgpu_nn_features(curs.data_ptr(), cw.data_ptr(), refs.data_ptr(), rss.data_ptr(), pos.data_ptr(),
                                  qps.data_ptr(), nmb, feats.data_ptr(), None) == 0
    torch.cuda.synchronize()
    del curs, refs, cw, rss, qps
    rel = np.rint(feats.cpu().numpy().reshape(nmb, 289, 2)[:, :, 1] * 256).astype(np.int64)
    p = np.arange(289)
    nrm = np.abs(p % 17 - 8) + np.abs(p // 17 - 8)
    return [np.lexsort((p, nrm, rel[i])) for i in range(nmb)]


def _score(mt, full, conf):
    m, f = L.metab_to_np(mt), L.metab_to_np(full)
    sel = conf >= 0.7
    hits = int((m["mv16"][sel, 0] == f["mv16"][sel, 0]).all(axis=1).sum())
    excess = int((m["sad16"][sel, 0] - f["sad16"][sel, 0]).sum())
    return hits, excess, int(sel.sum())


def _only(props, vecs, k):
    """Proposals whose candidate groups are the first k of vecs (the other
    groups repeat the first vector, so they add no search points)."""
    out = props.copy()
    for i in range(len(out)):
        v = list(vecs[i][:k]) + [vecs[i][0]] * (8 - k)
        out[i]["cx"] = [a for a, _ in v]
        out[i]["cy"] = [b for _, b in v]
        out[i]["kx"], out[i]["ky"] = v[0]
        out[i]["qx"], out[i]["qy"] = v[0]
    return out


def test_network_candidates_against_the_map_minima():
    tot = {}
    nsel = 0
    for sc, qp in ((content.scene_pan(640, 368, 81), 24), (content.scene_objects(640, 368, 82, nobj=8), 28),
                   (content.scene_blobs(640, 368, 83, nblob=150, vmax=24), 24),
                   (content.scene_pan(640, 368, 84, grain=1.0), 32)):
        w, h = sc.w, sc.h
        f0, f1 = sc.frame(0), sc.frame(1)
        an = G.Analyser([(w, h)], nsrc=1)
        an.set_proposer()
        an.put_ref(0, 0, f0)
        an.put_source(0, 0, f1)
        order = _maps(an, w, h, qp)              # against f0, before the step rotates the pool
        an.pframe([0], [qp], navail=[1], me_mode=G.ME_LEARNED)
        an.fetch()
        net = an.proposals(0)
        an.close()
        mapv = [[(2 * (q % 17 - 8), 2 * (q // 17 - 8)) for q in order[i][:8]] for i in range(len(net))]
        netv = [list(zip(net[i]["cx"].tolist(), net[i]["cy"].tolist())) for i in range(len(net))]
        for i in range(len(net)):
            assert list(zip(net[i]["kx"].tolist(), net[i]["ky"].tolist())) == mapv[i][:4], f"MB {i}: map minima"
        say(f"  ablation {w}x{h} qp {qp}")
        full = L.oracle_metab_exhaustive(f1, f0, None, w, h, 1)
        for k in (1, 8):
            for name, vecs in (("net", netv), ("map", mapv)):
                hits, excess, n = _score(L.oracle_metab_candidates(f1, f0, None, w, h, 1, _only(net, vecs, k)),
                                         full, net["conf"])
                t = tot.setdefault((name, k), [0, 0])
                t[0] += hits
                t[1] += excess
        nsel += n
    for k in (1, 8):
        say(f"  K = {k}: of {nsel} candidate-searched MBs the exhaustive vector is found with the network "
            f"{tot[('net', k)][0]}, with the map {tot[('map', k)][0]}; SAD given away {tot[('net', k)][1]} vs "
            f"{tot[('map', k)][1]}")
    assert nsel > 0
    for k in (1, 8):
        assert tot[("net", k)][1] < tot[("map", k)][1], f"K = {k}: the network gives away more SAD"
        assert tot[("net", k)][0] >= 0.99 * tot[("map", k)][0], f"K = {k}: the network finds far fewer vectors"
