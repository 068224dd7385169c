"""The sequence entry points with the reference's signatures and containers
(fdvamd_video_encode = fdv_video_encode, fdvamd_vtile_encode =
fdv_vtile_encode), key frames from the reference's own I-frame coder
registered through fdvamd_set_iframe_coder:

- in the diamond mode the whole stream -- headers, per-frame and per-band
  lengths, I- and P-frame blobs -- is byte-identical to the reference
  encoder's, for the 1280x720 eight-band workload (60 frames, key interval
  60, QP 24), the 320x180 single-tile workload (30 frames, QP 28) and a short
  stream with key frames every 5 frames;
- in the learned mode the tiled stream decodes with the reference decoder,
  band by band, to the frames the device encoder reconstructed.
"""
import numpy as np
import pytest

import orclib as L
from fdv_amd import content
from fdv_amd import gpu as G
from gpucases import say

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def ref_iframes():
    assert L.ref() is not None, "the reference build (oracle/_ref) is needed"
    enc, dec = L.ref_iframe_coder()
    G.set_iframe_coder(enc, dec)
    G.set_sequence_mode(G.ME_DIAMOND)
    yield
    G.set_sequence_mode(G.ME_DIAMOND)
    G.set_iframe_coder(0, 0)


def test_vtile_720p_60_frames_byte_identical_to_reference():
    sc = content.scene_pan(1280, 720, 2)
    frames = [sc.frame(i) for i in range(60)]
    say("  vtile 1280x720, 60 frames, keyint 60, QP 24, 8 bands: device encode")
    got = G.vtile_encode(frames, 1280, 720, 24, 60, 6, 8)
    say("  ... reference encode")
    want = L.ref_video_encode(frames, 1280, 720, 24, 60, band_mbrows=6)
    assert len(got) == len(want), f"{len(got)} B vs the reference's {len(want)} B"
    assert got == want
    dec = L.ref_video_decode(got, tiled=True)
    assert dec is not None and dec[0].shape[0] == 60


def test_video_320x180_30_frames_byte_identical_to_reference():
    sc = content.scene_objects(320, 180, 1)
    frames = [sc.frame(i) for i in range(30)]
    say("  video 320x192, 30 frames, QP 28")
    got = G.video_encode(frames, sc.w, sc.h, 28, 0)
    want = L.ref_video_encode(frames, sc.w, sc.h, 28, 0)
    assert got == want


def test_video_with_key_frames_every_5_frames():
    sc = content.scene_pan(96, 64, 23, nobj=3, smin=8, smax=32)
    frames = [sc.frame(i) for i in range(13)]
    say("  video 96x64, 13 frames, keyint 5")
    got = G.video_encode(frames, sc.w, sc.h, 26, 5)
    assert got == L.ref_video_encode(frames, sc.w, sc.h, 26, 5)
    got = G.vtile_encode(frames, sc.w, sc.h, 26, 5, 1, 2)
    assert got == L.ref_video_encode(frames, sc.w, sc.h, 26, 5, band_mbrows=1, nthreads=2)


def test_learned_vtile_decodes_to_the_device_reconstruction():
    sc = content.scene_pan(640, 368, 24)
    nfr = 6
    frames = [sc.frame(i) for i in range(nfr)]
    G.set_sequence_mode(G.ME_LEARNED)
    say("  vtile 640x368 learned, 6 frames, keyint 4")
    got = G.vtile_encode(frames, sc.w, sc.h, 24, 4, 6, 4)
    dec = L.ref_video_decode(got, tiled=True)
    assert dec is not None, "the reference decoder rejected the stream"
    frames_d, w, h = dec
    # the same bands through the drop-in batch encoder, reconstruction kept
    bands = [(y0, min(96, h - y0)) for y0 in range(0, h, 96)]
    enc = G.Encoder([(w, bh) for _, bh in bands], me_mode=G.ME_LEARNED)
    enc.set_proposer()
    tcs = [G.TabCache() for _ in bands]
    near_ok = 0

    def split(f):
        ys, cw = w * h, w // 2
        Y, U, V = f[:ys].reshape(h, w), f[ys:ys + ys // 4].reshape(h // 2, cw), f[ys + ys // 4:].reshape(h // 2, cw)
        return [np.concatenate([Y[y0:y0 + bh].ravel(), U[y0 // 2:(y0 + bh) // 2].ravel(),
                                V[y0 // 2:(y0 + bh) // 2].ravel()]) for y0, bh in bands]

    nref = 0
    for f in range(nfr):
        if f % 4 == 0:
            for b, sub in enumerate(split(frames_d[f])):
                enc.set_ref(b, 0, sub)
            nref = 1
            continue
        nav = [2 if nref >= 2 else 1] * len(bands)
        _, recs = enc.pframe_batch(split(frames[f]), nav, [24] * len(bands), tcs)
        for b, sub in enumerate(split(frames_d[f])):
            assert np.array_equal(recs[b], sub), f"frame {f} band {b}: decoded frame differs from the device recon"
            near_ok += 1
        nref = min(nref + 1, 2)
    enc.close()
    assert near_ok > 0
