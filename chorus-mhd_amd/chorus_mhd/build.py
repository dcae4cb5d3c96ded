"""Build libchorus_mhd.so (HIP kernels for gfx950 + the C-ABI) in-tree with hipcc.

Each translation unit is compiled separately because the FP64 kernels and the
host set-up arithmetic must not contract products and sums into FMAs (the
reference build emits none), while the mixed-precision kernels may.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                      # chorus-mhd_amd/
CSRC = os.path.join(ROOT, "csrc")
REPO = os.path.dirname(ROOT)
INCLUDE = os.path.join(REPO, "include")
LIB = os.path.join(ROOT, "libchorus_mhd.so")
OBJDIR = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-std=c++17", "-O3", "-fPIC", "-I" + INCLUDE, "-I" + CSRC, "-Wall", "-Wno-unused-function"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

# source -> extra flags
UNITS = {
    "chorus_fp64.hip": DEVICE + ["-ffp-contract=off"],
    "chorus_mixed.hip": DEVICE + ["-ffp-contract=fast"],
    "host_ops.cpp": ["--offload-arch=" + ARCH, "-ffp-contract=off"],
    "chorus_api.cpp": ["--offload-arch=" + ARCH, "-ffp-contract=off"],
    "tecplot.cpp": ["--offload-arch=" + ARCH, "-ffp-contract=off"],
}


def _compile(src: str, flags: list, objdir: str = OBJDIR) -> str:
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    if src.endswith(".cpp"):
        lang = ["-D__HIP_PLATFORM_AMD__"]
    cmd = [HIPCC] + COMMON + flags + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    warn = [l for l in r.stderr.splitlines() if "warning" in l and "packed-fp32-ops" not in l]
    if warn:
        sys.stderr.write("\n".join(warn[:20]) + "\n")
    return obj


def build(verbose: bool = False, out: str = LIB, defines: tuple = ()) -> str:
    """Compile every translation unit for gfx950 and link the shared library.
    ``defines``: extra -D tuning macros (variants go to another ``out`` path)."""
    if not shutil.which(HIPCC) and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found (%s)" % HIPCC)
    objdir = OBJDIR if out == LIB else os.path.join(OBJDIR, os.path.basename(out))
    os.makedirs(objdir, exist_ok=True)
    extra = ["-D" + d for d in defines]
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda kv: _compile(kv[0], kv[1] + extra, objdir), UNITS.items()))
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=LIB)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    args = ap.parse_args()
    build(verbose=True, out=os.path.abspath(args.out), defines=tuple(args.defines))
