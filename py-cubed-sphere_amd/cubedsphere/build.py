"""Build recipe of libcsadv.so: one hipcc call for gfx950, output in-tree next to _lib.py.

    python -m cubedsphere.build        (from py-cubed-sphere_amd/)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
REPO = os.path.dirname(PKG_ROOT)
SRC = os.path.join(PKG_ROOT, "csrc", "csadv.hip")
INCLUDE = os.path.join(REPO, "include")
OUT = os.path.join(HERE, "libcsadv.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def command(out=OUT, extra=()):
    return ["hipcc", *HIPCC_FLAGS, *extra, "-I", INCLUDE, "-I", os.path.join(PKG_ROOT, "csrc"), "-o", out, SRC]


def build_library(out=OUT, verbose=False):
    cmd = command(out)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building %s" % out)
    if verbose:
        sys.stdout.write(" ".join(cmd) + "\n")
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
