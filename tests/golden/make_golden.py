"""Writes tests/golden/ref_vectors.json from the reference codec compiled from
its own sources (oracle/_ref/libfdvref.so, built by `make -C oracle ref`, only
where the reference tree is present).

    python tests/golden/make_golden.py

The fixture holds expected outputs only (digests, tables, vectors and SADs);
the inputs come from vecgen.py and fdv_amd/content.py, both deterministic.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import blockrun  # noqa: E402


def main():
    out = {"source": "reference build (oracle/_ref/libfdvref.so) of yhbsh/fdv include/fdv.h",
           "blocks": blockrun.run_blocks("ref"),
           "chains": blockrun.run_chains("ref")}
    path = os.path.join(HERE, "ref_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=0, separators=(",", ":"))
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
