"""Build the native libraries in-tree.

* fdv_amd/libfdvgpu.so -- the HIP kernels for gfx950 and their C ABI
  (include/fdvgpu.h), plus the C11 host library (include/fdv_amd.h): entropy
  packing of P-frame payloads and the sequence drivers.
* oracle/ -- the parity oracle (test infrastructure), via oracle/Makefile.
* fdv_amd/learned/_scenes/ -- the scene library of the rate-distortion study,
  via fdv_amd/learned/Makefile (where the reference tree is present).

Run as `python -m fdv_amd.build` or through __graft_entry__.build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fdv_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfdvgpu.so")

HIP_SOURCES = ["fdvgpu_analyse.hip", "fdvgpu_post.hip", "fdvgpu_metab.hip", "fdvgpu_nn.hip", "fdvgpu_motion.hip", "fdvgpu_noalloc.hip", "fdvgpu_decode.hip", "fdvgpu_intra.hip", "fdvgpu_api.hip"]
C_SOURCES = ["fdv_pack.c", "fdv_video.c", "fdv_seq.c", "fdv_unpack.c", "fdv_decode.c", "fdv_stream.c", "fdv_image.c"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]
C_FLAGS = ["-std=c11", "-D_GNU_SOURCE", "-O2", "-fPIC", "-ffp-contract=off", "-Wall",
           "-Wno-unused-function", "-pthread", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), err))
    return err


def build_gpu(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        p = os.path.join(CSRC, src)
        if os.path.exists(p):
            jobs.append([HIPCC, *HIP_FLAGS, "-c", p, "-o", os.path.join(OBJ, src + ".o")])
    for src in C_SOURCES:
        p = os.path.join(HOST, src)
        if os.path.exists(p):
            jobs.append(["gcc", *C_FLAGS, "-c", p, "-o", os.path.join(OBJ, src + ".o")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    objs = [j[-1] for j in jobs]
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs, "-lm", "-pthread"])
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle() -> None:
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def build_scenes() -> None:
    """The scene library of the rate-distortion study (fdv_amd/learned/Makefile;
    a no-op where the reference tree is absent)."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "fdv_amd", "learned"), "scenes"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("scene library build failed:\n" + r.stdout + r.stderr)


def build_all(verbose: bool = False) -> None:
    build_gpu(verbose)
    build_oracle()             # after the library: the drop-in demonstration links it


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
