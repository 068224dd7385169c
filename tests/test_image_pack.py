"""The host half of the key-frame path (fdvgpu_image_pack, fdv_amd/host/fdv_image.c
and the container writer in fdv_pack.c), on the CPU: the oracle's decided
quadtree of a plane, serialised and packed by the product, must give the bytes
the oracle's own encoder gives -- which the reference's digests pin
(test_oracle_intra.py).  No device work."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import intraplanes as IP  # noqa: E402
import orclib as L  # noqa: E402
from fdv_amd import gpu as G  # noqa: E402


@pytest.mark.parametrize("case", IP.CASES, ids=lambda c: "%s-%dx%d-qp%d" % c[:4])
def test_pack_of_the_oracle_tree_matches_the_oracle_bytes(case):
    kind, w, h, qp, seed = case
    src = IP.intra_plane(kind, w, h, seed)
    want, _ = L.image_encode(L.oracle(), "orc_image_encode", src, qp)
    got = G.image_pack(L.image_tree(src, qp), w, h, qp)
    assert got.tobytes() == want.tobytes()
