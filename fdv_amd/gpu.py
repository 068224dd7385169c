"""ctypes binding of libfdvgpu.so (include/fdvgpu.h): the MI355X encode-analysis
path.  There is no fallback: if the library is missing or the GPU call fails,
this raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfdvgpu.so")

ME_DIAMOND, ME_EXHAUSTIVE, ME_LEARNED, ME_PARALLEL = 0, 1, 2, 3
NS = 16

MBINFO_DTYPE = np.dtype([("mode", "u1"), ("ref", "u1"), ("merge_idx", "u1"), ("sub", "u1"),
                         ("mv", "<i2", (4, 2))])
PROPOSAL_DTYPE = np.dtype([("cx", "i1", 8), ("cy", "i1", 8), ("kx", "i1", 4), ("ky", "i1", 4),
                           ("qx", "i1", (4, 2)), ("qy", "i1", (4, 2)), ("conf", "<f4"), ("pskip", "<f4")])
MOTION_DTYPE = np.dtype([("mv16", "<i2", (2, 2)), ("sad16", "<i4", 2), ("mv8", "<i2", (4, 2)),
                         ("sad8", "<i4", 4), ("flags", "<i4")])
MBCODE_DTYPE = np.dtype([("mode", "u1"), ("ref", "u1"), ("merge_idx", "u1"), ("ncand", "u1"), ("sub", "u1"),
                         ("ckind", "u1"), ("pad0", "u1"), ("pad1", "u1"), ("lflag", "u1", 4), ("cflag", "u1", 2),
                         ("pad2", "u1", 2), ("mvd", "<i2", (4, 2)), ("mv", "<i2", (4, 2)), ("lev", "<i2", 256),
                         ("clev", "<i2", 128)])
PROPOSER_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned", "proposer.f32")

_lib = None


def _signatures():
    """Every entry point this module (and the training script) calls, with its C
    signature: a handle passed without argtypes would be narrowed to a C int."""
    vp, i, f, d, sz = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_size_t
    ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
    u8pp = C.POINTER(C.POINTER(C.c_uint8))
    return {
        # include/fdvgpu.h
        "fdvgpu_open": (vp, [i, ip, ip, i]),
        "fdvgpu_close": (None, [vp]),
        "fdvgpu_last_error": (C.c_char_p, []),
        "fdvgpu_put_source": (i, [vp, i, i, vp]),
        "fdvgpu_put_ref": (i, [vp, i, i, vp]),
        "fdvgpu_set_given": (i, [vp, i, vp]),
        "fdvgpu_pframe": (i, [vp, ip, ip, dp, ip, ip, i]),
        "fdvgpu_fetch": (i, [vp]),
        "fdvgpu_streams": (i, [vp, i, u8pp, C.POINTER(sz)]),
        "fdvgpu_rewind": (i, [vp]),
        "fdvgpu_rewind_keep_motion": (i, [vp]),
        "fdvgpu_recon": (i, [vp, i, vp]),
        "fdvgpu_decisions": (i, [vp, i, vp]),
        "fdvgpu_set_graphs": (i, [vp, i]),
        "fdvgpu_set_overlap": (i, [vp, i]),
        "fdvgpu_get_overlap": (i, [vp]),
        "fdvgpu_get_tune": (i, [vp, vp]),
        "fdvgpu_set_tune": (i, [vp, vp]),
        "fdvgpu_autotune": (i, [vp, ip, ip, dp, ip, i, i]),
        "fdvgpu_nn_nparam": (i, []),
        "fdvgpu_set_proposer": (i, [vp, vp, i]),
        "fdvgpu_set_proposer_thresholds": (i, [vp, f, f]),
        "fdvgpu_proposals": (i, [vp, i, vp]),
        "fdvgpu_motion": (i, [vp, i, vp]),
        "fdvgpu_planes": (i, [vp, i, i, C.POINTER(vp), ip, C.POINTER(vp), ip]),
        "fdvgpu_nn_features": (i, [vp, vp, vp, vp, vp, vp, i, vp, vp]),
        "fdvgpu_nn_train_step": (i, [vp, vp, vp, vp, i, f, f, vp, vp, i, vp]),
        "fdvgpu_nn_adam": (i, [vp, vp, vp, vp, i, f, i, i, vp]),
        "fdvgpu_timings": (i, [vp, C.POINTER(C.c_float)]),
        "fdvgpu_submit_slots": (i, [vp, i, ip, ip, dp, ip, ip, i]),
        "fdvgpu_collect": (i, [vp, i]),
        "fdvgpu_collected_streams": (i, [vp, i, i, vp, vp]),
        "fdvgpu_slot_timings": (i, [vp, i, C.POINTER(C.c_float)]),
        "fdvgpu_decode": (i, [vp, vp, ip, ip]),
        "fdvgpu_motion_search": (i, [vp, vp, vp, i, i, i, i, i, vp, f, vp, vp]),
        "fdvgpu_motion_search_noalloc": (i, [vp, vp, vp, i, i, i, i, i, vp, f, vp, vp]),
        "fdvgpu_itree_bytes": (sz, [i, i]),
        "fdvgpu_image_batch": (i, [i, vp, vp]),
        "fdvgpu_image_pack": (sz, [vp, i, i, i, vp, sz]),
        # include/fdv_amd.h
        "fdvamd_pack_pframe": (sz, [i, vp, vp, vp, vp, sz]),
        "fdvamd_encoder_open": (vp, [i, vp, vp, i]),
        "fdvamd_encoder_close": (None, [vp]),
        "fdvamd_last_error": (C.c_char_p, []),
        "fdvamd_set_ref": (i, [vp, i, i, vp]),
        "fdvamd_set_ref_planes": (i, [vp, i, i, vp, vp, vp]),
        "fdvamd_pframe_encode": (sz, [vp, vp, vp, vp, i, i, i, i, vp, sz, vp, vp, vp, i, vp, d]),
        "fdvamd_pframe_batch": (i, [vp] * 11),
        "fdvamd_pframe_submit": (i, [vp] * 6),
        "fdvamd_pframe_collect": (i, [vp] * 5),
        "fdvamd_set_proposer": (i, [vp, vp, i]),
        "fdvamd_set_proposer_thresholds": (i, [vp, f, f]),
        "fdvamd_rewind": (i, [vp]),
        "fdvamd_rewind_keep_motion": (i, [vp]),
        "fdvamd_set_iframe_coder": (i, [vp, vp]),
        "fdvamd_unpack_pframe": (i, [vp, sz, sz, vp, ip, C.POINTER(vp), C.POINTER(sz)]),
        "fdvamd_parse_pframe": (i, [vp, C.POINTER(sz), i, i, vp]),
        "fdvamd_decoder_open": (vp, [i, ip, ip]),
        "fdvamd_decoder_close": (None, [vp]),
        "fdvamd_decoder_set_ref": (i, [vp, i, vp]),
        "fdvamd_pframe_decode_batch": (i, [vp, vp, vp, vp, vp, vp]),
        "fdvamd_video_decode": (i, [vp, sz, vp, ip, ip, ip]),
        "fdvamd_vtile_decode": (i, [vp, sz, vp, ip, ip, ip, i]),
        "fdvamd_set_sequence_mode": (i, [i]),
        "fdvamd_set_sequence_proposer": (i, [vp, i]),
        "fdvamd_video_encode": (sz, [vp, i, i, i, i, i, vp, sz]),
        "fdvamd_vtile_encode": (sz, [vp, i, i, i, i, i, i, i, vp, sz]),
        "fdvamd_image_encode": (sz, [vp, i, i, i, i, vp, sz, vp]),
        "fdvamd_image_encode_batch": (i, [i, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "fdvamd_iframe_encode": (sz, [vp, vp, vp, i, i, i, vp, sz, vp, vp, vp]),
        "fdvamd_image_decode": (i, [vp, sz, vp, i, ip, ip]),
        "fdvamd_iframe_decode": (i, [vp, sz, vp, vp, vp, i, i]),
    }


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("fdv_amd/libfdvgpu.so is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _signatures().items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ints(v):
    return (C.c_int * len(v))(*[int(x) for x in v])


class GpuError(RuntimeError):
    pass


class Analyser:
    """Batched P-frame analysis for a set of sequences (streams or bands)."""

    def __init__(self, sizes, nsrc: int = 1):
        self.L = lib()
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.nseq = len(self.sizes)
        ws = _ints([w for w, _ in self.sizes])
        hs = _ints([h for _, h in self.sizes])
        self.h = self.L.fdvgpu_open(self.nseq, ws, hs, int(nsrc))
        if not self.h:
            raise GpuError("fdvgpu_open: " + self.L.fdvgpu_last_error().decode())

    def _ck(self, rc, what):
        if rc != 0:
            raise GpuError(what + ": " + self.L.fdvgpu_last_error().decode())

    def close(self):
        if self.h:
            self.L.fdvgpu_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def frame_bytes(self, seq):
        w, h = self.sizes[seq]
        return w * h + 2 * (w // 2) * (h // 2)

    def put_source(self, seq, slot, frame: np.ndarray):
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        assert f.size == self.frame_bytes(seq)
        self._ck(self.L.fdvgpu_put_source(self.h, seq, slot, f.ctypes.data), "put_source")

    def put_ref(self, seq, which, frame: np.ndarray):
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        assert f.size == self.frame_bytes(seq)
        self._ck(self.L.fdvgpu_put_ref(self.h, seq, which, f.ctypes.data), "put_ref")

    def set_given(self, seq, dec):
        if dec is None:
            self._ck(self.L.fdvgpu_set_given(self.h, seq, None), "set_given")
            return
        d = np.ascontiguousarray(dec, dtype=MBINFO_DTYPE)
        self._ck(self.L.fdvgpu_set_given(self.h, seq, d.ctypes.data), "set_given")

    def pframe(self, src_slot, qp, lscale=None, navail=None, force_skip=None, me_mode=ME_DIAMOND):
        n = self.nseq
        lscale = [1.0] * n if lscale is None else lscale
        navail = [2] * n if navail is None else navail
        force_skip = [0] * n if force_skip is None else force_skip
        ls = (C.c_double * n)(*[float(x) for x in lscale])
        self._ck(self.L.fdvgpu_pframe(self.h, _ints(src_slot), _ints(qp), ls, _ints(navail),
                                      _ints(force_skip), int(me_mode)), "pframe")

    def fetch(self):
        self._ck(self.L.fdvgpu_fetch(self.h), "fetch")

    def submit(self, slot, src_slot, qp, lscale=None, navail=None, force_skip=None, me_mode=ME_DIAMOND):
        """Queue a frame step from the source slots into staging slot 0 or 1
        without waiting (fdvgpu_submit_slots); collect(slot) before reusing it."""
        n = self.nseq
        lscale = [1.0] * n if lscale is None else lscale
        navail = [2] * n if navail is None else navail
        force_skip = [0] * n if force_skip is None else force_skip
        ls = (C.c_double * n)(*[float(x) for x in lscale])
        self._ck(self.L.fdvgpu_submit_slots(self.h, int(slot), _ints(src_slot), _ints(qp), ls, _ints(navail),
                                            _ints(force_skip), int(me_mode)), "submit_slots")

    def collect(self, slot):
        """Wait for the step staged in `slot` and check its flags."""
        self._ck(self.L.fdvgpu_collect(self.h, int(slot)), "collect")

    def collected_streams(self, slot, seq):
        sv = (C.POINTER(C.c_uint8) * NS)()
        sn = (C.c_size_t * NS)()
        self._ck(self.L.fdvgpu_collected_streams(self.h, int(slot), seq, sv, sn), "collected_streams")
        return [bytes(C.string_at(sv[k], sn[k])) if sn[k] else b"" for k in range(NS)]

    def slot_timings(self, slot):
        ms = (C.c_float * 4)()
        self._ck(self.L.fdvgpu_slot_timings(self.h, int(slot), ms), "slot_timings")
        return list(ms)

    def rewind(self, keep_motion=False):
        """Undo the last frame step's reference rotation (fdvgpu_rewind); with
        keep_motion the next step reuses its motion tables."""
        fn = self.L.fdvgpu_rewind_keep_motion if keep_motion else self.L.fdvgpu_rewind
        self._ck(fn(self.h), "rewind")

    def streams(self, seq):
        sv = (C.POINTER(C.c_uint8) * NS)()
        sn = (C.c_size_t * NS)()
        self._ck(self.L.fdvgpu_streams(self.h, seq, sv, sn), "streams")
        return [bytes(C.string_at(sv[k], sn[k])) if sn[k] else b"" for k in range(NS)]

    def recon(self, seq):
        out = np.zeros(self.frame_bytes(seq), dtype=np.uint8)
        self._ck(self.L.fdvgpu_recon(self.h, seq, out.ctypes.data), "recon")
        return out

    def decisions(self, seq):
        w, h = self.sizes[seq]
        out = np.zeros((w // 16) * (h // 16), dtype=MBINFO_DTYPE)
        self._ck(self.L.fdvgpu_decisions(self.h, seq, out.ctypes.data), "decisions")
        return out

    def set_proposer(self, weights=None):
        """Install the learned proposer's weights (default: the trained ones
        shipped with the package)."""
        w = load_proposer() if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        self._ck(self.L.fdvgpu_set_proposer(self.h, w.ctypes.data, int(w.size)), "set_proposer")

    def set_proposer_thresholds(self, conf=0.7, skip=2.0):
        """Confidence threshold of the candidate search and SKIP threshold of the
        pruning (above 1: never; the default)."""
        self._ck(self.L.fdvgpu_set_proposer_thresholds(self.h, float(conf), float(skip)), "set_proposer_thresholds")

    def set_overlap(self, mode):
        """The block-parallel decision stage's structure: 0 separate launches
        (costs, replays, coding kernels); 1 the costs, the replays, then the
        coding workers on one stream; 2 the replays beside the costs and the
        coding workers beside the replays (the default where the device holds
        them side by side; True means 2); 3 the replays beside the costs, then
        the coding workers; 4 as 2, with more workers joining once the replays
        are done (the default where 2 is available).  A structure set here is
        kept by autotune()."""
        mode = 2 if mode is True else int(mode)
        self._ck(self.L.fdvgpu_set_overlap(self.h, mode), "set_overlap")

    def overlap(self):
        """The decision structure in use (0 to 4; see set_overlap)."""
        v = self.L.fdvgpu_get_overlap(self.h)
        if v < 0:
            raise GpuError("get_overlap: " + self.L.fdvgpu_last_error().decode())
        return v

    def set_graphs(self, on: bool):
        self._ck(self.L.fdvgpu_set_graphs(self.h, 1 if on else 0), "set_graphs")

    def tune(self, shapes=None):
        """Get (or set) the launch shapes: [serialiser waves per workgroup,
        deblock workgroup size, proposer workgroups per CU, decision coding
        workers per CU]."""
        if shapes is not None:
            self._ck(self.L.fdvgpu_set_tune(self.h, _ints(shapes)), "set_tune")
        out = (C.c_int * 4)()
        self._ck(self.L.fdvgpu_get_tune(self.h, out), "get_tune")
        return list(out)

    def autotune(self, src_slot, qp, lscale=None, navail=None, me_mode=ME_DIAMOND, reps=3):
        n = self.nseq
        ls = (C.c_double * n)(*([1.0] * n if lscale is None else lscale))
        self._ck(self.L.fdvgpu_autotune(self.h, _ints(src_slot), _ints(qp), ls,
                                        _ints([2] * n if navail is None else navail), int(me_mode), int(reps)),
                 "autotune")
        return self.tune()

    def proposals(self, seq):
        w, h = self.sizes[seq]
        out = np.zeros((w // 16) * (h // 16), dtype=PROPOSAL_DTYPE)
        self._ck(self.L.fdvgpu_proposals(self.h, seq, out.ctypes.data), "proposals")
        return out

    def motion(self, seq):
        w, h = self.sizes[seq]
        out = np.zeros((w // 16) * (h // 16), dtype=MOTION_DTYPE)
        self._ck(self.L.fdvgpu_motion(self.h, seq, out.ctypes.data), "motion")
        return out

    def planes(self, seq, slot):
        """Device pointers: (source luma, its stride, nearest reference luma, its stride)."""
        cy, ry = C.c_void_p(), C.c_void_p()
        cs, rs = C.c_int(), C.c_int()
        self._ck(self.L.fdvgpu_planes(self.h, seq, slot, C.byref(cy), C.byref(cs), C.byref(ry), C.byref(rs)),
                 "planes")
        return cy.value, cs.value, ry.value, rs.value

    def timings(self):
        ms = (C.c_float * 4)()
        self._ck(self.L.fdvgpu_timings(self.h, ms), "timings")
        return list(ms)


TABCACHE_BYTES = 2 * 256 + 2 * 16 * 256 + 4 + 4 * 16 * 256 + 4


class TabCache:
    """The entropy state carried between P-frames (fdvamd_tabcache)."""

    def __init__(self):
        self.buf = np.zeros(TABCACHE_BYTES, dtype=np.uint8)


def pack_pframe(qp: int, streams, tc: TabCache | None = None) -> bytes:
    """Host entropy packing of one P-frame payload (fdvamd_pack_pframe)."""
    L = lib()
    arrs = [np.frombuffer(s, dtype=np.uint8) if len(s) else np.zeros(1, dtype=np.uint8) for s in streams]
    sv = (C.c_void_p * NS)(*[a.ctypes.data for a in arrs])
    sn = (C.c_size_t * NS)(*[len(s) for s in streams])
    cap = sum(len(s) for s in streams) * 3 + 65536
    out = np.zeros(cap, dtype=np.uint8)
    n = L.fdvamd_pack_pframe(int(qp), C.addressof(sv), C.addressof(sn),
                             tc.buf.ctypes.data if tc is not None else None, out.ctypes.data, cap)
    if n == 0:
        raise GpuError("fdvamd_pack_pframe failed")
    return out[:n].tobytes()


class Encoder:
    """The drop-in P-frame encoder (include/fdv_amd.h): device analysis plus host
    entropy packing, with the near/far reference pool kept on the device and
    rotated after every P-frame (fdv_video_encode's pool, fdv.h:6842-6846)."""

    def __init__(self, sizes, me_mode: int = ME_DIAMOND):
        self.L = lib()
        L = self.L
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        n = len(self.sizes)
        self.h = L.fdvamd_encoder_open(n, _ints([w for w, _ in self.sizes]),
                                       _ints([h for _, h in self.sizes]), int(me_mode))
        if not self.h:
            raise GpuError("fdvamd_encoder_open: " + L.fdvamd_last_error().decode())

    def close(self):
        if self.h:
            self.L.fdvamd_encoder_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ref(self, seq, which, frame):
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        if self.L.fdvamd_set_ref(self.h, seq, which, f.ctypes.data) != 0:
            raise GpuError("fdvamd_set_ref: " + self.L.fdvamd_last_error().decode())

    def set_proposer(self, weights=None):
        w = load_proposer() if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        if self.L.fdvamd_set_proposer(self.h, w.ctypes.data, int(w.size)) != 0:
            raise GpuError("fdvamd_set_proposer: " + self.L.fdvamd_last_error().decode())

    def set_proposer_thresholds(self, conf=0.7, skip=2.0):
        if self.L.fdvamd_set_proposer_thresholds(self.h, float(conf), float(skip)) != 0:
            raise GpuError("fdvamd_set_proposer_thresholds: " + self.L.fdvamd_last_error().decode())

    def set_ref_planes(self, seq, which, frame):
        """Install a reference from three separate tight planes (the form the
        reference's encoders hold their reconstruction in)."""
        w, h = self.sizes[seq]
        ys, cs = w * h, (w // 2) * (h // 2)
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        y, u, v = f[:ys].copy(), f[ys:ys + cs].copy(), f[ys + cs:].copy()
        if self.L.fdvamd_set_ref_planes(self.h, seq, which, y.ctypes.data, u.ctypes.data, v.ctypes.data) != 0:
            raise GpuError("fdvamd_set_ref_planes: " + self.L.fdvamd_last_error().decode())

    def rewind(self, keep_motion=False):
        fn = self.L.fdvamd_rewind_keep_motion if keep_motion else self.L.fdvamd_rewind
        if fn(self.h) != 0:
            raise GpuError("fdvamd_rewind: " + self.L.fdvamd_last_error().decode())

    def pframe_encode(self, frame, navail, qp, tc: TabCache | None = None, lscale=1.0, force_skip=0):
        """One P-frame of sequence 0: (payload bytes, recon I420)."""
        w, h = self.sizes[0]
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        ys, cs = w * h, (w // 2) * (h // 2)
        cap = ys * 3 + 65536
        out = np.zeros(cap, np.uint8)
        rec = np.zeros(ys + 2 * cs, np.uint8)
        n = self.L.fdvamd_pframe_encode(self.h, f.ctypes.data, f.ctypes.data + ys, f.ctypes.data + ys + cs,
                                        w, h, int(navail), int(qp), out.ctypes.data, cap,
                                        rec.ctypes.data, rec.ctypes.data + ys, rec.ctypes.data + ys + cs,
                                        int(force_skip), tc.buf.ctypes.data if tc is not None else None,
                                        float(lscale))
        if n == 0:
            raise GpuError("fdvamd_pframe_encode: " + self.L.fdvamd_last_error().decode())
        return out[:n].tobytes(), rec

    def pframe_batch(self, frames, navail, qp, tcs=None, lscale=None, force_skip=None, recon=True):
        """One P-frame of every sequence in one device pass: lists of payloads and
        recons (None without `recon`: the reconstruction then stays on the device)."""
        n = len(self.sizes)
        fr = [np.ascontiguousarray(f, dtype=np.uint8) for f in frames]
        caps = [w * h * 3 + 65536 for w, h in self.sizes]
        outs = [np.empty(c, np.uint8) for c in caps]
        recs = [np.empty(w * h + 2 * (w // 2) * (h // 2), np.uint8) for w, h in self.sizes] if recon else None
        P = C.c_void_p * n
        lens = (C.c_size_t * n)()
        ls = (C.c_double * n)(*([1.0] * n if lscale is None else lscale))
        tcp = P(*[t.buf.ctypes.data for t in tcs]) if tcs is not None else None
        # every argument array is bound to a name so that it outlives the call
        fp = P(*[f.ctypes.data for f in fr])
        nav = _ints(navail)
        qps = _ints(qp)
        fsk = _ints(force_skip if force_skip is not None else [0] * n)
        op = P(*[o.ctypes.data for o in outs])
        cp = (C.c_size_t * n)(*caps)
        rp = P(*[r.ctypes.data for r in recs]) if recs is not None else None
        rc = self.L.fdvamd_pframe_batch(
            self.h, C.addressof(fp), C.addressof(nav), C.addressof(qps), C.addressof(ls),
            C.addressof(fsk), C.addressof(tcp) if tcp is not None else None, C.addressof(op),
            C.addressof(cp), C.addressof(lens), C.addressof(rp) if rp is not None else None)
        if rc != 0:
            raise GpuError("fdvamd_pframe_batch: " + self.L.fdvamd_last_error().decode())
        return [outs[s][:lens[s]].tobytes() for s in range(n)], recs

    def pframe_submit(self, frames, navail, qp, lscale=None, force_skip=None):
        """Queue one P-frame of every sequence (fixed-QP pipelining, at most two in
        flight); returns without waiting for the device."""
        n = len(self.sizes)
        fr = [np.ascontiguousarray(f, dtype=np.uint8) for f in frames]
        P = C.c_void_p * n
        fp = P(*[f.ctypes.data for f in fr])
        nav = _ints(navail)
        qps = _ints(qp)
        ls = (C.c_double * n)(*([1.0] * n if lscale is None else lscale))
        fsk = _ints(force_skip) if force_skip is not None else None
        rc = self.L.fdvamd_pframe_submit(self.h, C.addressof(fp), C.addressof(nav), C.addressof(qps),
                                         C.addressof(ls), C.addressof(fsk) if fsk is not None else None)
        if rc != 0:
            raise GpuError("fdvamd_pframe_submit: " + self.L.fdvamd_last_error().decode())

    def pframe_collect(self, tcs=None):
        """Payloads of the oldest submitted frame, packed while the next one is
        analysed."""
        n = len(self.sizes)
        caps = [w * h * 3 + 65536 for w, h in self.sizes]
        outs = [np.empty(c, np.uint8) for c in caps]
        P = C.c_void_p * n
        op = P(*[o.ctypes.data for o in outs])
        cp = (C.c_size_t * n)(*caps)
        lens = (C.c_size_t * n)()
        tcp = P(*[t.buf.ctypes.data for t in tcs]) if tcs is not None else None
        rc = self.L.fdvamd_pframe_collect(self.h, C.addressof(tcp) if tcp is not None else None, C.addressof(op),
                                          C.addressof(cp), C.addressof(lens))
        if rc != 0:
            raise GpuError("fdvamd_pframe_collect: " + self.L.fdvamd_last_error().decode())
        return [outs[s][:lens[s]].tobytes() for s in range(n)]


def motion_search(cur_y: int, ref0_y: int, ref1_y: int, ref_stride: int, w: int, h: int, navail: int, qp: int,
                  table: int, weights: int = 0, conf: float = 0.7, stream: int = 0, noalloc: bool = False):
    """fdvgpu_motion_search (or, with `noalloc`, fdvgpu_motion_search_noalloc)
    on device pointers (ints): the motion table of a whole frame into `table`
    (MOTION_DTYPE entries, raster order)."""
    fn = lib().fdvgpu_motion_search_noalloc if noalloc else lib().fdvgpu_motion_search
    rc = fn(cur_y, ref0_y, ref1_y or None, ref_stride, w, h, navail, qp, weights or None,
                                    float(conf), table, stream or None)
    if rc != 0:
        raise GpuError(lib().fdvgpu_last_error().decode())


def set_iframe_coder(encode_fn: int, decode_fn: int = 0):
    """Register the key-frame coder of the sequence entry points: C function
    pointers (ints) with the fdvamd_iframe_encode_fn / _decode_fn contracts."""
    lib().fdvamd_set_iframe_coder(encode_fn or None, decode_fn or None)


def set_sequence_mode(me_mode: int, weights=None):
    """Search mode of video_encode / vtile_encode (and, for the learned mode,
    the proposer's weights: default the shipped ones)."""
    L = lib()
    if me_mode == ME_LEARNED:
        w = load_proposer() if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        if L.fdvamd_set_sequence_proposer(w.ctypes.data, int(w.size)) != 0:
            raise GpuError(L.fdvamd_last_error().decode())
    if L.fdvamd_set_sequence_mode(int(me_mode)) != 0:
        raise GpuError(L.fdvamd_last_error().decode())


def _frame_ptrs(frames):
    fr = [np.ascontiguousarray(f, dtype=np.uint8) for f in frames]
    return fr, (C.c_void_p * len(fr))(*[f.ctypes.data for f in fr])


def video_encode(frames, w: int, h: int, qp: int, keyint: int) -> bytes:
    """fdvamd_video_encode: the reference's fdv_video_encode container."""
    fr, fp = _frame_ptrs(frames)
    cap = len(fr) * (w * h * 3 + 8192) + 64
    out = np.empty(cap, np.uint8)
    n = lib().fdvamd_video_encode(C.addressof(fp), len(fr), w, h, qp, keyint, out.ctypes.data, cap)
    if n == 0:
        raise GpuError("fdvamd_video_encode: " + lib().fdvamd_last_error().decode())
    return out[:n].tobytes()


def vtile_encode(frames, w: int, h: int, qp: int, keyint: int, band_mbrows: int, nthreads: int = 8) -> bytes:
    """fdvamd_vtile_encode: the reference's fdv_vtile_encode container."""
    fr, fp = _frame_ptrs(frames)
    cap = len(fr) * (w * h * 3 + 8192) + 65536
    out = np.empty(cap, np.uint8)
    n = lib().fdvamd_vtile_encode(C.addressof(fp), len(fr), w, h, qp, keyint, band_mbrows, nthreads,
                                  out.ctypes.data, cap)
    if n == 0:
        raise GpuError("fdvamd_vtile_encode: " + lib().fdvamd_last_error().decode())
    return out[:n].tobytes()


def unpack_pframe(payload: bytes, nmb: int, tc: TabCache | None = None):
    """fdvamd_unpack_pframe: (qp, the 16 symbol streams) of a P-frame payload;
    raises GpuError for a payload the reference decoder rejects."""
    L = lib()
    buf = np.frombuffer(bytes(payload), dtype=np.uint8).copy() if payload else np.zeros(1, np.uint8)
    qp = C.c_int()
    syms = C.c_void_p()
    sn = (C.c_size_t * NS)()
    rc = L.fdvamd_unpack_pframe(buf.ctypes.data, len(payload), nmb, tc.buf.ctypes.data if tc is not None else None,
                                C.byref(qp), C.byref(syms), sn)
    if rc != 0:
        raise GpuError("fdvamd_unpack_pframe: malformed payload")
    out, off = [], 0
    for k in range(NS):
        out.append(C.string_at(syms.value + off, sn[k]) if sn[k] else b"")
        off += sn[k]
    C.CDLL(None).free(C.c_void_p(syms.value))
    return qp.value, out


def parse_pframe(streams, w: int, h: int) -> np.ndarray:
    """fdvamd_parse_pframe: the MB codes (MBCODE_DTYPE, raster order) of a
    P-frame's 16 symbol streams; raises GpuError where the reference decoder
    would reject them."""
    allb = b"".join(streams)
    buf = np.frombuffer(allb, dtype=np.uint8).copy() if allb else np.zeros(1, np.uint8)
    sn = (C.c_size_t * NS)(*[len(s) for s in streams])
    codes = np.zeros((w // 16) * (h // 16), dtype=MBCODE_DTYPE)
    if lib().fdvamd_parse_pframe(buf.ctypes.data, sn, w, h, codes.ctypes.data) != 0:
        raise GpuError("fdvamd_parse_pframe: malformed streams")
    return codes


def video_decode(stream: bytes, tiled: bool = False, nthreads: int = 8):
    """fdvamd_video_decode / fdvamd_vtile_decode: (frames [n, w*h*3/2], w, h)."""
    b = bytes(stream[:8])
    if tiled:
        w, h, n = b[0] | b[1] << 8, b[2] | b[3] << 8, b[4] | b[5] << 8
    else:
        n, w, h = b[0] | b[1] << 8, b[2] | b[3] << 8, b[4] | b[5] << 8
    data = np.frombuffer(bytes(stream), dtype=np.uint8).copy()
    out = np.zeros(max(n * (w * h * 3 // 2), 1), np.uint8)
    nf, ww, hh = C.c_int(), C.c_int(), C.c_int()
    L = lib()
    if tiled:
        rc = L.fdvamd_vtile_decode(data.ctypes.data, len(data), out.ctypes.data, C.byref(nf), C.byref(ww),
                                   C.byref(hh), nthreads)
    else:
        rc = L.fdvamd_video_decode(data.ctypes.data, len(data), out.ctypes.data, C.byref(nf), C.byref(ww),
                                   C.byref(hh))
    if rc != 0:
        raise GpuError("fdvamd_video_decode: " + L.fdvamd_last_error().decode())
    return out[:n * (w * h * 3 // 2)].reshape(n, -1), w, h


class Decoder:
    """The P-frame decoder (fdvamd_decoder): payloads unpacked and parsed on the
    host, frames reconstructed on the device against its reference pool."""

    def __init__(self, sizes):
        self.L = lib()
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        self.h = self.L.fdvamd_decoder_open(len(self.sizes), _ints([w for w, _ in self.sizes]),
                                            _ints([h for _, h in self.sizes]))
        if not self.h:
            raise GpuError("fdvamd_decoder_open: " + self.L.fdvamd_last_error().decode())

    def close(self):
        if self.h:
            self.L.fdvamd_decoder_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ref(self, seq, frame):
        f = np.ascontiguousarray(frame, dtype=np.uint8)
        if self.L.fdvamd_decoder_set_ref(self.h, seq, f.ctypes.data) != 0:
            raise GpuError("fdvamd_decoder_set_ref: " + self.L.fdvamd_last_error().decode())

    def pframe(self, payloads, navail, tcs=None):
        """One P-frame of every sequence: the decoded frames (tight I420)."""
        n = len(self.sizes)
        bufs = [np.frombuffer(bytes(p), dtype=np.uint8).copy() if len(p) else np.zeros(1, np.uint8) for p in payloads]
        outs = [np.zeros(w * h * 3 // 2, np.uint8) for w, h in self.sizes]
        P = C.c_void_p * n
        ip_ = P(*[b.ctypes.data for b in bufs])
        ln = (C.c_size_t * n)(*[len(p) for p in payloads])
        nav = _ints(navail)
        tcp = P(*[t.buf.ctypes.data for t in tcs]) if tcs is not None else None
        op = P(*[o.ctypes.data for o in outs])
        rc = self.L.fdvamd_pframe_decode_batch(self.h, C.addressof(ip_), C.addressof(ln), C.addressof(nav),
                                               C.addressof(tcp) if tcp is not None else None, C.addressof(op))
        if rc != 0:
            raise GpuError("fdvamd_pframe_decode_batch: " + self.L.fdvamd_last_error().decode())
        return outs


def nn_nparam() -> int:
    return int(lib().fdvgpu_nn_nparam())


def load_proposer(path: str = PROPOSER_WEIGHTS) -> np.ndarray:
    if not os.path.exists(path):
        raise GpuError("the learned proposer's weights are missing: " + path)
    w = np.fromfile(path, dtype="<f4")
    if w.size != nn_nparam():
        raise GpuError("proposer weights: expected %d floats, found %d" % (nn_nparam(), w.size))
    return w


# ---- key frames (fdvamd_image_encode_batch, fdvgpu_image_batch, fdvgpu_image_pack)

class _ITreeOut(C.Structure):
    _fields_ = [("rec", C.c_void_p), ("lev", C.c_void_p), ("rflag", C.c_void_p), ("cmode", C.c_void_p),
                ("tree", C.c_void_p)]


class _IPlaneIn(C.Structure):
    _fields_ = [("src", C.c_void_p), ("w", C.c_int), ("h", C.c_int), ("stride", C.c_int), ("qp", C.c_int)]


def _tree_arrays(w, h):
    L = lib()
    return {"rec": np.zeros((h, w), np.uint8), "lev": np.zeros((h, w), np.int16),
            "rflag": np.zeros(((h + 7) // 8, (w + 7) // 8), np.uint8), "cmode": np.zeros((h // 4, w // 4), np.uint8),
            "tree": np.zeros(int(L.fdvgpu_itree_bytes(w, h)), np.uint8)}


def _tree_struct(t):
    return _ITreeOut(t["rec"].ctypes.data, t["lev"].ctypes.data, t["rflag"].ctypes.data, t["cmode"].ctypes.data,
                     t["tree"].ctypes.data)


def image_trees(planes, qps):
    """The device decisions of fdvgpu_image_batch for tight planes: per plane a
    dict of rec (deblocked), lev, rflag, cmode and the tree planes split<d> /
    nmode<d>."""
    L = lib()
    planes = [np.ascontiguousarray(p, dtype=np.uint8) for p in planes]
    ins = (_IPlaneIn * len(planes))(*[_IPlaneIn(p.ctypes.data, p.shape[1], p.shape[0], p.shape[1], int(q))
                                      for p, q in zip(planes, qps)])
    outs = [_tree_arrays(p.shape[1], p.shape[0]) for p in planes]
    st = (_ITreeOut * len(planes))(*[_tree_struct(t) for t in outs])
    if L.fdvgpu_image_batch(len(planes), ins, st) != 0:
        raise GpuError(L.fdvgpu_last_error().decode())
    for p, t in zip(planes, outs):
        h, w = p.shape
        o = 0
        for d in range(4):
            n = 64 >> d
            th, tw = -(-h // n), -(-w // n)
            t["split%d" % d] = t["tree"][o:o + tw * th].reshape(th, tw)
            t["nmode%d" % d] = t["tree"][o + tw * th:o + 2 * tw * th].reshape(th, tw)
            o += 2 * tw * th
    return outs


def image_pack(t, w, h, qp):
    """fdvgpu_image_pack: a decided plane (dict as image_trees returns, or the
    same planes from elsewhere) serialised into the image container."""
    L = lib()
    tree = np.concatenate([np.concatenate([np.ascontiguousarray(t["split%d" % d]).ravel(),
                                           np.ascontiguousarray(t["nmode%d" % d]).ravel()]) for d in range(4)])
    a = {"rec": np.zeros((h, w), np.uint8), "lev": np.ascontiguousarray(t["lev"], np.int16),
         "rflag": np.ascontiguousarray(t["rflag"], np.uint8), "cmode": np.ascontiguousarray(t["cmode"], np.uint8),
         "tree": np.ascontiguousarray(tree, np.uint8)}
    cap = w * h * 6 + 8192
    out = np.zeros(cap, np.uint8)
    st = _tree_struct(a)
    n = L.fdvgpu_image_pack(C.byref(st), w, h, qp, out.ctypes.data, cap)
    if not n:
        raise GpuError(L.fdvamd_last_error().decode())
    return out[:n].copy()


def image_encode_batch(planes, qps):
    """fdvamd_image_encode_batch on tight planes: [(container bytes, recon)]."""
    L = lib()
    planes = [np.ascontiguousarray(p, dtype=np.uint8) for p in planes]
    n = len(planes)
    caps = [p.size * 6 + 8192 for p in planes]
    outs = [np.zeros(c, np.uint8) for c in caps]
    recs = [np.zeros_like(p) for p in planes]
    P = lambda arrs: (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
    ws = _ints([p.shape[1] for p in planes])
    hs = _ints([p.shape[0] for p in planes])
    lens = (C.c_size_t * n)()
    rc = L.fdvamd_image_encode_batch(n, P(planes), ws, hs, ws, _ints(qps), P(outs), (C.c_size_t * n)(*caps),
                                     P(recs), lens)
    if rc != 0:
        raise GpuError(L.fdvamd_last_error().decode())
    return [(o[:lens[i]].copy(), r) for i, (o, r) in enumerate(zip(outs, recs))]
