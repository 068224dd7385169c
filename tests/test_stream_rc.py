"""The average-bitrate control (fdvamd_rc_*, a host restatement of the
reference's fdv_rc_*, fdv.h:7880-8040): its arithmetic decides each frame's
quantizer and lambda scale, so the streaming encoder's output matches the
reference's only if every step matches bit for bit.  Driven here side by side
with the reference's own functions (oracle/_ref) over seeded call sequences --
key frames and P-frames, frames far over and under their allocation, held
frames, adopted operating points, splits across the quantizer range -- the two
states must stay byte-identical.  A planted mistake in a copy of the state
(one step of the gain) must be caught."""
import ctypes as C
import random

import pytest

import orclib as L
from fdv_amd import gpu as G
from streamlib import Rc, Streaming


@pytest.fixture(scope="module")
def libs():
    if L.ref() is None:
        pytest.skip("the reference build (oracle/_ref) is not present")
    return Streaming(G.lib(), "fdvamd_"), Streaming(L.ref(), "fdv_")


def run(s, seed, steps=400, plant=False):
    rnd = random.Random(seed)
    rc = Rc()
    w, h = rnd.choice([(320, 192), (1280, 720), (1920, 1088)])
    s.f("rc_init")(C.byref(rc), rnd.choice([0, 150000, 1000000, 8000000]), rnd.choice([0, 25, 30, 60]), w, h,
                   rnd.choice([0, 1, 30, 60, 255]), rnd.choice([0.0, 0.5, 1.0, 3.0]), rnd.choice([0, 12]),
                   rnd.choice([0, 40, 51]))
    trace = [bytes(rc)]
    planted = False
    for k in range(steps):
        op = rnd.random()
        intra = 1 if rnd.random() < 0.1 else 0
        if op < 0.55:
            s.f("rc_pick")(C.byref(rc), intra)
            target = max(rc.last_target, 1.0)
            size = int(target / 8 * rnd.choice([0.02, 0.3, 0.9, 1.0, 1.6, 4.0, 60.0]) * rnd.uniform(0.8, 1.2))
            s.f("rc_update")(C.byref(rc), intra, max(size, 0))
            if plant and not planted and k >= steps // 2:
                rc.qp_p += 1.0 / 1024
                planted = True
        elif op < 0.7:
            s.f("rc_skip")(C.byref(rc), rnd.randrange(0, 5000))
        elif op < 0.85:
            s.f("rc_adopt")(C.byref(rc), intra, rnd.uniform(0.0, 60.0))
        else:
            qp, ls = C.c_int(), C.c_double()
            s.f("rc_split")(C.byref(rc), rnd.uniform(-5.0, 60.0), C.byref(qp), C.byref(ls))
            trace.append(bytes(C.c_int(qp.value)) + bytes(ls))
        trace.append(bytes(rc))
    return trace


@pytest.mark.parametrize("seed", range(12))
def test_rate_control_matches_the_reference_bit_for_bit(libs, seed):
    ours, ref = libs
    assert run(ours, seed) == run(ref, seed)


def test_a_planted_drift_is_caught(libs):
    ours, ref = libs
    assert run(ours, 3, plant=True) != run(ref, 3)
