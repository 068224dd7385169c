"""ctypes binding of the streaming surface, for either library: the
reference's fdv_* (oracle/_ref) or this library's fdvamd_*, which share
signatures and struct layouts (fdv_info / fdvamd_info, fdv_rc / fdvamd_rc)."""
import ctypes as C

import numpy as np


class Info(C.Structure):
    _fields_ = [("kind", C.c_int), ("version", C.c_int), ("fps", C.c_int), ("w", C.c_int), ("h", C.c_int),
                ("nframes", C.c_int), ("qp", C.c_int), ("keyint", C.c_int), ("nbands", C.c_int),
                ("payload", C.c_void_p), ("payload_len", C.c_size_t)]


class Rc(C.Structure):
    _fields_ = [("bitrate", C.c_int), ("fps", C.c_int), ("bufbits", C.c_double), ("qpmin", C.c_int),
                ("qpmax", C.c_int), ("keyint", C.c_int), ("qp_p", C.c_double), ("ip_off", C.c_double),
                ("last_target", C.c_double), ("fill", C.c_double), ("used_qp", C.c_int), ("lscale", C.c_double)]


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


class Streaming:
    def __init__(self, lib, prefix):
        self.L, self.p = lib, prefix
        vp, i, sz, d = C.c_void_p, C.c_int, C.c_size_t, C.c_double
        sig = {"enc_open": (vp, [i] * 4), "enc_open_tiled": (vp, [i] * 6), "enc_frame": (i, [vp, vp]),
               "enc_repeat": (i, [vp]), "enc_set_bitrate": (i, [vp, i, i, d]), "enc_count": (i, [vp]),
               "enc_bytes": (sz, [vp]), "enc_finish": (vp, [vp, i, C.POINTER(C.c_size_t)]),
               "enc_close": (None, [vp]), "wrap": (sz, [vp, sz, i, i, vp, sz]),
               "read": (i, [vp, sz, C.POINTER(Info)]), "decode": (i, [C.POINTER(Info), vp, i]),
               "dec_open": (vp, [C.POINTER(Info)]), "dec_next": (i, [vp, vp]), "dec_seek": (i, [vp, i]),
               "dec_pos": (i, [vp]), "dec_frame_size": (sz, [vp]), "dec_close": (None, [vp]),
               "rc_init": (None, [C.POINTER(Rc), i, i, i, i, i, d, i, i]), "rc_pick": (i, [C.POINTER(Rc), i]),
               "rc_update": (None, [C.POINTER(Rc), i, sz]), "rc_skip": (None, [C.POINTER(Rc), sz]),
               "rc_adopt": (None, [C.POINTER(Rc), i, d]),
               "rc_split": (None, [C.POINTER(Rc), d, C.POINTER(C.c_int), C.POINTER(C.c_double)])}
        for name, (res, args) in sig.items():
            fn = getattr(lib, prefix + name)
            fn.restype, fn.argtypes = res, args

    def f(self, name):
        return getattr(self.L, self.p + name)

    def encode(self, frames, w, h, qp, keyint, tiled=None, bitrate=0, hold=(), fps=30):
        """The streaming encoder over `frames`, a held frame before each index in
        `hold`; the finished file's bytes."""
        e = self.f("enc_open_tiled")(w, h, qp, keyint, *tiled) if tiled else self.f("enc_open")(w, h, qp, keyint)
        assert e, "enc_open failed"
        if bitrate:
            assert self.f("enc_set_bitrate")(e, bitrate, fps, 1.0) == 0
        fr = [np.ascontiguousarray(x, dtype=np.uint8) for x in frames]
        for k, x in enumerate(fr):
            if k in hold:
                assert self.f("enc_repeat")(e) == 0
            assert self.f("enc_frame")(e, x.ctypes.data) == 0, f"frame {k}"
        n = C.c_size_t()
        buf = self.f("enc_finish")(e, fps, C.byref(n))
        assert buf, "enc_finish failed"
        data = C.string_at(buf, n.value)
        _libc.free(buf)
        return data

    def read(self, data):
        arr = np.frombuffer(data, dtype=np.uint8).copy()
        info = Info()
        rc = self.f("read")(arr.ctypes.data, len(arr), C.byref(info))
        return rc, info, arr

    def decoder(self, data):
        rc, info, arr = self.read(data)
        assert rc == 0
        d = self.f("dec_open")(C.byref(info))
        assert d, "dec_open failed"
        return Decoder(self, d, info, arr)


class Decoder:
    def __init__(self, s, d, info, arr):
        self.s, self.d, self.info, self.arr = s, d, info, arr     # info and arr stay alive with the decoder
        self.fs = s.f("dec_frame_size")(d)

    def next(self):
        out = np.zeros(self.fs, np.uint8)
        rc = self.s.f("dec_next")(self.d, out.ctypes.data)
        return rc, out

    def seek(self, frame):
        return self.s.f("dec_seek")(self.d, frame)

    def pos(self):
        return self.s.f("dec_pos")(self.d)

    def close(self):
        self.s.f("dec_close")(self.d)
