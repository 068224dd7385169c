"""The key-frame decoder (fdvamd_image_decode / fdvamd_iframe_decode, host
C in fdv_amd/host/fdv_unpack.c), on the CPU:

- every golden plane's container, as the reference's digests pin it (the
  oracle writes the same bytes), decodes to the encoder's reconstruction --
  and, where the reference build is present, to what the reference's
  fdv_image_decode returns, also for random planes and whole key frames;
- input the reference rejects is rejected: truncated containers, a bad
  container mode, a plane whose header disagrees with the frame, corrupt
  region flags.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import intraplanes as IP  # noqa: E402
import orclib as L  # noqa: E402
from fdv_amd import gpu as G  # noqa: E402


def _decode(lib, fn, data, w, h):
    out = np.full((h, w), 7, np.uint8)
    buf = np.frombuffer(bytes(data), np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    wo, ho = C.c_int(0), C.c_int(0)
    rc = getattr(lib, fn)(buf.ctypes.data, len(data), out.ctypes.data, w, C.byref(wo), C.byref(ho))
    return rc, out


@pytest.mark.parametrize("case", IP.CASES, ids=lambda c: "%s-%dx%d-qp%d" % c[:4])
def test_golden_containers_decode_to_the_reconstruction(case):
    kind, w, h, qp, seed = case
    data, rec = L.image_encode(L.oracle(), "orc_image_encode", IP.intra_plane(kind, w, h, seed), qp)
    rc, out = _decode(G.lib(), "fdvamd_image_decode", data.tobytes(), w, h)
    assert rc == 0
    assert (out == rec).all()


needs_ref = pytest.mark.skipif(L.ref() is None, reason="reference build (oracle/_ref) not present")


@needs_ref
def test_agrees_with_the_reference_decoder():
    ref = L.ref()
    ref.fdv_image_decode.restype = C.c_int
    ref.fdv_image_decode.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    r = np.random.default_rng(9)
    for i in range(20):
        w, h = int(r.integers(1, 50)) * 4, int(r.integers(1, 40)) * 4
        qp = int(r.integers(0, 52))
        data, rec = L.image_encode(L.oracle(), "orc_image_encode",
                                   IP.intra_plane(IP.KINDS[i % len(IP.KINDS)], w, h, 50 + i), qp)
        a = _decode(G.lib(), "fdvamd_image_decode", data.tobytes(), w, h)
        b = _decode(ref, "fdv_image_decode", data.tobytes(), w, h)
        assert a[0] == b[0] == 0
        assert (a[1] == b[1]).all() and (a[1] == rec).all()
        # corruptions: both accept or both reject, and agree on what they accept
        for cut in (3, 7, len(data) // 2, len(data) - 1):
            x = _decode(G.lib(), "fdvamd_image_decode", data.tobytes()[:cut], w, h)
            y = _decode(ref, "fdv_image_decode", data.tobytes()[:cut], w, h)
            assert (x[0] == 0) == (y[0] == 0), (w, h, qp, cut)
            if x[0] == 0:
                assert (x[1] == y[1]).all()
        bad = bytearray(data.tobytes())
        bad[5] = 3                                        # no such container mode
        assert _decode(G.lib(), "fdvamd_image_decode", bytes(bad), w, h)[0] == -1
        assert _decode(ref, "fdv_image_decode", bytes(bad), w, h)[0] == -1


def test_key_frame_round_trip():
    w, h = 160, 96
    y = IP.intra_plane("mixed", w, h, 1)
    u = IP.intra_plane("ramp", w // 2, h // 2, 2)
    v = IP.intra_plane("sine", w // 2, h // 2, 3)
    cap = w * h * 8 + 65536
    out = np.zeros(cap, np.uint8)
    ry, ru, rv = np.zeros_like(y), np.zeros_like(u), np.zeros_like(v)
    n = L.oracle().orc_iframe_encode(y.ctypes.data, u.ctypes.data, v.ctypes.data, w, h, 27, out.ctypes.data, cap,
                                     ry.ctypes.data, ru.ctypes.data, rv.ctypes.data)
    assert n > 0
    dy, du, dv = np.zeros_like(y), np.zeros_like(u), np.zeros_like(v)
    lib = G.lib()
    assert lib.fdvamd_iframe_decode(out.ctypes.data, n, dy.ctypes.data, du.ctypes.data, dv.ctypes.data, w, h) == 0
    assert (dy == ry).all() and (du == ru).all() and (dv == rv).all()
    # a frame of another size is refused before anything is written
    assert lib.fdvamd_iframe_decode(out.ctypes.data, n, dy.ctypes.data, du.ctypes.data, dv.ctypes.data, w, h + 16) == -1
    assert lib.fdvamd_iframe_decode(out.ctypes.data, n - 1, dy.ctypes.data, du.ctypes.data, dv.ctypes.data, w, h) == -1
