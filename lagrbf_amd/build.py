"""Build liblagrbf.so (HIP, gfx950) in-tree with hipcc."""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblagrbf.so")
SOURCES = ["eval_blk_wc2.hip", "eval_blk_phs3.hip", "eval_blk_gauss.hip", "eval_blk_other.hip",
           "context.hip", "kernels.hip", "eval_cell.hip", "eval_blk.hip", "serve.hip", "patch_chol.hip", "adaptive.hip", "aug.hip", "chol_mfma.hip", "chol_tiled.hip", "solve_llt.hip", "aug_lean.hip", "trsv_tiles.hip", "noalloc.hip", "devmem.cpp", "options.cpp", "io.cpp"]
ARCH = os.environ.get("LAGRBF_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=8, incremental=False):
    """Compile every source to an object (in parallel), then link the shared library.
    incremental: keep objects newer than their source and every header."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    hdr_time = max(os.path.getmtime(os.path.join(CSRC, f)) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h")))
    common = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-Wno-unused-result", "-I", os.path.join(HERE, "..", "include")]
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(
                hdr_time, os.path.getmtime(os.path.join(CSRC, src))):
            objs.append(obj)
            continue
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [hipcc()] + common + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    for src, p in procs:
        out, _ = p.communicate()
        msg = out.decode(errors="replace")
        if p.returncode != 0:
            sys.stderr.write(msg)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose and msg.strip():
            sys.stderr.write(msg)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="-i" in sys.argv))
