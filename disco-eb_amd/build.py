"""Build libdiscoeb_amd.so (gfx950) in-tree with hipcc.

    python disco-eb_amd/build.py          # or __graft_entry__.build()
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdiscoeb_amd.so")
SOURCES = ["deb_perturb.hip", "deb_adjoint.hip", "deb_background.hip", "deb_misc.hip", "deb_cmb.hip", "deb_lss.hip", "deb_noalloc.hip", "deb_sources.hip", "deb_capi.cpp"]
ARCH = os.environ.get("DEB_OFFLOAD_ARCH", "gfx950")


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def _flags():
    # -ffp-contract=on: a multiply-add is fused only within one source expression, so every
    # instantiation of a kernel template (threads per unit) rounds the same expressions the same
    # way; the default lets the backend fuse across statements depending on each one's dataflow.
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-ffp-contract=on", "-munsafe-fp-atomics", "-Wno-unused-result", "-I", CSRC]


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()

    def compile_one(src):
        out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *_flags(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
        if verbose and r.stderr.strip():
            print(r.stderr[-3000:], file=sys.stderr)
        return out

    with ThreadPoolExecutor(max_workers=5) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
