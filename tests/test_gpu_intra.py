"""Key frames on the device (fdvamd_image_encode_batch / fdvamd_image_encode /
fdvamd_iframe_encode, the rate-distortion quadtree in k_intra_unit and
k_intra_emit) against the oracle (orc_image_encode, pinned to the reference
build by test_oracle_intra.py) and the reference's own digests:

- the 16 golden planes in one device batch: container bytes and deblocked
  reconstruction equal the reference's digests;
- seeded planes of random size and QP, a whole 1280x720 plane and its
  chroma, against the oracle byte for byte, with the decided tree compared
  unit by unit where they differ (to name the first unit that does);
- whole key frames (three planes) against the oracle's iframe_encode;
- the sequence encoders with the device key frames (no coder registered):
  fdvamd_vtile_encode at 1280x720 in 8 bands and fdvamd_video_encode with key
  frames every 3 frames are byte-identical to the reference's streams.
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import intraplanes as IP  # noqa: E402
import orclib as L  # noqa: E402
from fdv_amd import content  # noqa: E402
from fdv_amd import gpu as G  # noqa: E402
from gpucases import say  # noqa: E402

pytestmark = pytest.mark.gpu

with open(os.path.join(HERE, "golden", "intra_vectors.json")) as f:
    FIX = json.load(f)["cases"]


def _first_difference(src, qp, got_bytes):
    """Where the device's decisions first differ from the oracle's (unit, what)."""
    h, w = src.shape
    dev = G.image_trees([src], [qp])[0]
    orc = L.image_tree(src, qp)
    for uy in range(0, h, 64):
        for ux in range(0, w, 64):
            sl = np.s_[uy // 4:(uy + 64) // 4, ux // 4:(ux + 64) // 4]
            if not (dev["cmode"][sl] == orc["cmode"][sl]).all():
                return "unit (%d, %d): 4x4 mode map differs" % (ux, uy)
            ls = np.s_[uy:uy + 64, ux:ux + 64]
            if not (dev["lev"][ls] == orc["lev"][ls]).all():
                return "unit (%d, %d): levels differ" % (ux, uy)
    return "trees agree; container differs (%d bytes)" % len(got_bytes)


def test_golden_planes_in_one_batch_match_the_reference_digests():
    planes = [IP.intra_plane(*[c["case"][i] for i in (0, 1, 2, 4)]) for c in FIX]
    qps = [c["case"][3] for c in FIX]
    say("  %d planes in one device batch" % len(planes))
    res = G.image_encode_batch(planes, qps)
    for c, (out, rec) in zip(FIX, res):
        assert hashlib.sha256(out.tobytes()).hexdigest() == c["bytes_sha256"], c["case"]
        assert hashlib.sha256(rec.tobytes()).hexdigest() == c["rec_sha256"], c["case"]


def test_random_planes_match_the_oracle():
    r = np.random.default_rng(77)
    planes, qps = [], []
    for i in range(12):
        w, h = int(r.integers(1, 60)) * 4, int(r.integers(1, 40)) * 4
        planes.append(IP.intra_plane(IP.KINDS[i % len(IP.KINDS)], w, h, 300 + i))
        qps.append(int(r.integers(0, 52)))
    say("  12 random planes in one device batch")
    res = G.image_encode_batch(planes, qps)
    for src, qp, (out, rec) in zip(planes, qps, res):
        want, wrec = L.image_encode(L.oracle(), "orc_image_encode", src, qp)
        if out.tobytes() != want.tobytes():
            pytest.fail("%dx%d qp %d: %s" % (src.shape[1], src.shape[0], qp, _first_difference(src, qp, out)))
        assert (rec == wrec).all()


@pytest.mark.parametrize("qp", [8, 24, 40])
def test_720p_frame_planes_match_the_oracle(qp):
    sc = content.scene_pan(1280, 720, 3)
    f = sc.frame(0)
    y, u, v = L.split_i420(f, 1280, 720)
    planes = [y.reshape(720, 1280), u.reshape(360, 640), v.reshape(360, 640)]
    say("  1280x720 key frame planes, qp %d" % qp)
    res = G.image_encode_batch(planes, [qp] * 3)
    for src, (out, rec) in zip(planes, res):
        want, wrec = L.image_encode(L.oracle(), "orc_image_encode", src, qp)
        if out.tobytes() != want.tobytes():
            pytest.fail("%dx%d: %s" % (src.shape[1], src.shape[0], _first_difference(src, qp, out)))
        assert (rec == wrec).all()


def test_single_plane_entry_point():
    src = IP.intra_plane("mixed", 200, 120, 5)
    lib = G.lib()
    cap = 200 * 120 * 4 + 8192
    out = np.zeros(cap, np.uint8)
    rec = np.zeros_like(src)
    n = lib.fdvamd_image_encode(src.ctypes.data, 200, 120, 200, 30, out.ctypes.data, cap, rec.ctypes.data)
    want, wrec = L.image_encode(L.oracle(), "orc_image_encode", src, 30)
    assert out[:n].tobytes() == want.tobytes()
    assert (rec == wrec).all()
    # the reference's argument checks: sizes not multiples of 4 code nothing
    assert lib.fdvamd_image_encode(src.ctypes.data, 198, 120, 200, 30, out.ctypes.data, cap, None) == 0


def test_key_frame_matches_the_oracle():
    sc = content.scene_objects(320, 192, 4)
    f = sc.frame(0)
    y, u, v = L.split_i420(f, 320, 192)
    lib = G.lib()
    cap = 320 * 192 * 3 + 65536
    outs = []
    for fn, lb in (("fdvamd_iframe_encode", lib), ("orc_iframe_encode", L.oracle())):
        out = np.zeros(cap, np.uint8)
        rec = np.zeros(320 * 192 * 3 // 2, np.uint8)
        ry, ru, rv = L.split_i420(rec, 320, 192)
        n = getattr(lb, fn)(y.ctypes.data, u.ctypes.data, v.ctypes.data, 320, 192, 26, out.ctypes.data, cap,
                            ry.ctypes.data, ru.ctypes.data, rv.ctypes.data)
        assert n > 0
        outs.append((out[:n].tobytes(), rec))
    assert outs[0][0] == outs[1][0]
    assert (outs[0][1] == outs[1][1]).all()


@pytest.fixture
def device_key_frames():
    """No key-frame encoder registered (the device codes them); the
    reference's decoder registered for the decode checks."""
    assert L.ref() is not None, "the reference build (oracle/_ref) is needed"
    _, dec = L.ref_iframe_coder()
    G.set_iframe_coder(0, dec)
    G.set_sequence_mode(G.ME_DIAMOND)
    yield
    G.set_iframe_coder(0, 0)


def test_vtile_with_device_key_frames_byte_identical_to_reference(device_key_frames):
    sc = content.scene_pan(1280, 720, 2)
    frames = [sc.frame(i) for i in range(8)]
    say("  vtile 1280x720, 8 frames, keyint 4, QP 24, 8 bands, device key frames")
    got = G.vtile_encode(frames, 1280, 720, 24, 4, 6, 8)
    say("  ... reference encode")
    want = L.ref_video_encode(frames, 1280, 720, 24, 4, band_mbrows=6)
    assert len(got) == len(want), f"{len(got)} B vs the reference's {len(want)} B"
    assert got == want


def test_video_with_device_key_frames_byte_identical_to_reference(device_key_frames):
    sc = content.scene_objects(320, 192, 1)
    frames = [sc.frame(i) for i in range(7)]
    say("  video 320x192, 7 frames, keyint 3, QP 30, device key frames")
    got = G.video_encode(frames, 320, 192, 30, 3)
    want = L.ref_video_encode(frames, 320, 192, 30, 3)
    assert got == want


def test_consecutive_batches_of_different_layouts():
    """Each call's units wait only for words written in that call: a batch laid
    out differently from the one before reuses its device memory, and must not
    read the earlier contents as published edges."""
    big = [IP.intra_plane("mixed", 320, 256, 1), IP.intra_plane("sine", 168, 144, 2)]
    small = [IP.intra_plane("noise", 132, 128, 301), IP.intra_plane("mixed", 100, 36, 4),
             IP.intra_plane("checker", 208, 56, 302)]
    qps = [12, 12, 50]
    want = [L.image_encode(L.oracle(), "orc_image_encode", p, q) for p, q in zip(small, qps)]
    for rep in range(3):
        G.image_encode_batch(big, [20, 30])
        got = G.image_encode_batch(small, qps)
        for (a, ra), (b, rb) in zip(got, want):
            assert a.tobytes() == b.tobytes(), rep
            assert (ra == rb).all(), rep


def test_streams_round_trip_with_no_coder_registered():
    """Nothing of the reference in the loop: key frames coded on the device and
    decoded on the host by the built-in coders, P-frames both ways on the
    device; the streams decode to what the reference decoder makes of them."""
    G.set_iframe_coder(0, 0)
    G.set_sequence_mode(G.ME_DIAMOND)
    sc = content.scene_pan(640, 368, 5)
    frames = [sc.frame(i) for i in range(6)]
    say("  vtile 640x368, 6 frames, keyint 3, QP 26, 4 bands, built-in key-frame coders")
    tiled = G.vtile_encode(frames, 640, 368, 26, 3, 6, 4)
    ours = G.video_decode(tiled, tiled=True)
    theirs = L.ref_video_decode(tiled, tiled=True)
    assert theirs is not None
    assert (ours[0] == theirs[0]).all()
    say("  video 320x192, 5 frames, keyint 2, QP 33")
    sc2 = content.scene_objects(320, 192, 6)
    single = G.video_encode([sc2.frame(i) for i in range(5)], 320, 192, 33, 2)
    ours = G.video_decode(single)
    theirs = L.ref_video_decode(single)
    assert theirs is not None
    assert (ours[0] == theirs[0]).all()
