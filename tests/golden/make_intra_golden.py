"""Writes tests/golden/intra_vectors.json from the reference codec compiled
from its own sources (oracle/_ref/libfdvref.so, built by `make -C oracle ref`
where the reference tree is present): for every plane of intraplanes.CASES the
digest of fdv_image_encode's bytes and of its deblocked reconstruction, and
the container mode it chose.

    python tests/golden/make_intra_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import intraplanes as IP  # noqa: E402
import orclib as L  # noqa: E402


def main():
    ref = L.ref()
    if ref is None:
        raise SystemExit("oracle/_ref is not built")
    cases = []
    for kind, w, h, qp, seed in IP.CASES:
        src = IP.intra_plane(kind, w, h, seed)
        out, rec = L.image_encode(ref, "fdv_image_encode", src, qp)
        cases.append({"case": [kind, w, h, qp, seed], "len": int(len(out)), "container": int(out[5]),
                      "bytes_sha256": hashlib.sha256(out.tobytes()).hexdigest(),
                      "rec_sha256": hashlib.sha256(rec.tobytes()).hexdigest()})
    path = os.path.join(HERE, "intra_vectors.json")
    with open(path, "w") as f:
        json.dump({"source": "reference build (oracle/_ref/libfdvref.so), fdv_image_encode", "cases": cases},
                  f, indent=1)
    print("wrote", path, "containers", sorted({c["container"] for c in cases}))


if __name__ == "__main__":
    main()
