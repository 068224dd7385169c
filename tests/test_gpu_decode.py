"""P-frame decoding on the device (fdvgpu_decode behind fdvamd_decoder):
payloads unpacked and parsed on the host, frames reconstructed and deblocked
on the GPU, every sequence of a step in one pass.

- Payloads of the device encoder in each search mode, over chains of P-frames
  in a mixed-size batch, decode to the encoder's reconstruction, which the
  reference decoder (pframe_decode, fdv.h:6492) also produces.
- Intra-heavy content (fast blobs, a scene cut, grain) and QP 0 / 51 exercise
  the intra pass and the 8x8 transforms.
- fdvamd_video_decode / fdvamd_vtile_decode decode the reference encoder's own
  streams (key frames through the reference's I-frame decoder) to exactly the
  frames fdv_video_decode / fdv_vtile_decode produce.
"""
import numpy as np
import pytest

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import say

pytestmark = pytest.mark.gpu


def _decode_chain(scenes, nfr, qps, mode):
    sizes = [(s.w, s.h) for s in scenes]
    n = len(scenes)
    frames = [[s.frame(i) for i in range(nfr)] for s in scenes]
    enc = G.Encoder(sizes, me_mode=mode)
    if mode == G.ME_LEARNED:
        enc.set_proposer()
    dec = G.Decoder(sizes)
    near = [frames[k][0] for k in range(n)]
    far = [None] * n
    for k in range(n):
        enc.set_ref(k, 0, near[k])
        dec.set_ref(k, near[k])
    tce = [G.TabCache() for _ in range(n)]
    tcd = [G.TabCache() for _ in range(n)]
    tcr = [L.TabCache() for _ in range(n)]
    for f in range(1, nfr):
        nav = [1 if f == 1 else 2] * n
        pays, recs = enc.pframe_batch([frames[k][f] for k in range(n)], nav, qps, tce)
        say(f"  decode mode {mode} frame {f}: {n} sequences")
        outs = dec.pframe(pays, nav, tcd)
        for k in range(n):
            w, h = sizes[k]
            assert np.array_equal(outs[k], recs[k]), f"seq {k} ({w}x{h}) frame {f}: device decode != encoder recon"
            ref = L.ref_pframe_decode(pays[k], near[k], far[k], w, h, nav[k], tcr[k])
            assert ref is not None and np.array_equal(ref, recs[k]), f"seq {k} frame {f}: reference decoder"
            far[k], near[k] = near[k], recs[k]
    enc.close()
    dec.close()


@pytest.mark.parametrize("mode", [G.ME_DIAMOND, G.ME_EXHAUSTIVE, G.ME_PARALLEL, G.ME_LEARNED])
def test_decode_matches_encoder_and_reference(mode):
    scenes = [content.scene_pan(160, 96, 51, nobj=4, smin=16, smax=48),
              content.scene_objects(128, 48, 52, nobj=3, smin=8, smax=32, vmax=10, grain=2.0),
              content.scene_blobs(96, 64, 53, nblob=25, vmax=12, grain=3.0, cut=3)]
    _decode_chain(scenes, 5, [24, 30, 20], mode)


@pytest.mark.parametrize("qp", [0, 51])
def test_decode_qp_extremes(qp):
    _decode_chain([content.scene_pan(96, 64, 60 + qp, grain=4.0, nobj=2, smin=8, smax=24)], 4, [qp], G.ME_DIAMOND)


def test_decode_intra_heavy_720p_bands():
    sc = content.scene_blobs(1280, 720, 5)
    layout = [(y0, min(96, 720 - y0)) for y0 in range(0, 720, 96)]

    class Band:
        def __init__(self, y0, bh):
            self.w, self.h, self.y0 = 1280, bh, y0

        def frame(self, i):
            y, u, v = sc.planes(i)
            return np.concatenate([y[self.y0:self.y0 + self.h].ravel(), u[self.y0 // 2:(self.y0 + self.h) // 2].ravel(),
                                   v[self.y0 // 2:(self.y0 + self.h) // 2].ravel()])

    _decode_chain([Band(y0, bh) for y0, bh in layout], 3, [24] * len(layout), G.ME_PARALLEL)


@pytest.fixture
def ref_iframes():
    assert L.ref() is not None, "the reference build (oracle/_ref) is needed"
    enc, dec = L.ref_iframe_coder()
    G.set_iframe_coder(enc, dec)
    yield
    G.set_iframe_coder(0, 0)


def test_video_decode_of_reference_streams(ref_iframes):
    sc = content.scene_pan(320, 192, 71, nobj=5, smin=16, smax=64)
    frames = [sc.frame(i) for i in range(12)]
    stream = L.ref_video_encode(frames, sc.w, sc.h, 26, 5)
    say("  fdvamd_video_decode of a reference stream, 12 frames, keyint 5")
    got, w, h = G.video_decode(stream)
    want, _, _ = L.ref_video_decode(stream)
    assert (w, h) == (sc.w, sc.h)
    assert np.array_equal(got, want)


def test_vtile_decode_720p_eight_bands(ref_iframes):
    sc = content.scene_pan(1280, 720, 2)
    frames = [sc.frame(i) for i in range(10)]
    stream = L.ref_video_encode(frames, 1280, 720, 24, 60, band_mbrows=6)
    say("  fdvamd_vtile_decode of the reference's 1280x720 eight-band stream, 10 frames")
    got, w, h = G.video_decode(stream, tiled=True)
    want, _, _ = L.ref_video_decode(stream, tiled=True)
    assert np.array_equal(got, want)
