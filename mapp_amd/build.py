"""Build recipe of the MI355X DMD library (libmapp_dmd.so, in-tree).

    python -m mapp_amd.build          # or __graft_entry__.build()

hipcc cross-compiles for gfx950 here; the .so travels with the tree.
Flags: -ffp-contract=off keeps every product/sum separately rounded unless a
kernel asks for fma() explicitly (the neighbour list must be bit-identical to
the reference's); packed FP32 ops are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "lib", "libmapp_dmd.so")
SOURCES = ["capi.hip"]


def hipcc_cmd(out, sources):
    return ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17",
            "-fPIC", "-shared", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-o", out] + sources


def build(verbose=True):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = hipcc_cmd(tmp, srcs)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines()
            if "packed-fp32-ops" not in l]
    if msgs and verbose:
        print("\n".join(msgs))
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
