"""Build recipe of libcsadv.so for gfx950, output in-tree next to _lib.py.

    python -m cubedsphere.build        (from py-cubed-sphere_amd/)

Every translation unit of csrc/ compiles to an object of its own (in parallel: the fused step
kernels are split by precision and reconstruction), then hipcc links the shared library.
The objects go to py-cubed-sphere_amd/_build/ (not part of the sources).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
REPO = os.path.dirname(PKG_ROOT)
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(REPO, "include")
OUT = os.path.join(HERE, "libcsadv.so")
OBJ_DIR = os.path.join(PKG_ROOT, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep the compiler from emitting packed FP32 arithmetic
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
# host-only planning code (no device code)
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC"]


def sources():
    """Device translation units (.hip) and the host-only planning unit (.cpp)."""
    names = sorted(os.listdir(CSRC))
    return [os.path.join(CSRC, n) for n in names if n.endswith(".hip") or n.endswith(".cpp")]


# The FP32 fused step kernels fuse a multiply and an add only within one source expression.
# hipcc's default fuses them across statements wherever they share a basic block, so the
# rounding follows each kernel's code layout: the one-solve and the multi-solve kernels (a solve
# alone and in a batch), built from the same march, must give the same bits, and the FP32 march's
# branch around its last row's row stage is enough to make them fuse differently.  The same VALU
# count (the products the default fuses are mostly needed on their own as well), and these
# launches are bound by issue.  The FP64 kernels keep the default: they are bound by latency,
# where the longer chains of separate multiplies and adds measured 1.5-2% slower.
STEP_FLAGS = ["-ffp-contract=on"]


def _compile_cmd(src, obj):
    inc = ["-I", INCLUDE, "-I", CSRC]
    if src.endswith(".cpp"):
        return ["hipcc", *CXX_FLAGS, *inc, "-c", "-o", obj, src]
    extra = STEP_FLAGS if os.path.basename(src).startswith("step_f32_") else []
    return ["hipcc", *HIPCC_FLAGS, *extra, *inc, "-c", "-o", obj, src]


def command(out=OUT):
    """The link command (objects from compile_commands())."""
    objs = [_obj(s) for s in sources()]
    return ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out, *objs]


def _obj(src):
    return os.path.join(OBJ_DIR, os.path.basename(src) + ".o")


def compile_commands():
    return [_compile_cmd(s, _obj(s)) for s in sources()]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    return cmd, res


def build_library(out=OUT, verbose=False, jobs=None):
    os.makedirs(OBJ_DIR, exist_ok=True)
    cmds = compile_commands()
    jobs = jobs or min(len(cmds), max(1, min(16, os.cpu_count() or 1)))
    failed = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd, res in ex.map(_run, cmds):
            if res.returncode != 0:
                failed.append((cmd, res))
            elif verbose:
                sys.stdout.write(" ".join(cmd) + "\n")
    if failed:
        for cmd, res in failed:
            sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError("hipcc failed on %d translation unit(s)" % len(failed))
    tmp = out + ".tmp"
    cmd, res = _run(command(tmp))
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("linking %s failed" % out)
    os.replace(tmp, out)
    if verbose:
        sys.stdout.write(" ".join(command(out)) + "\n")
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
