"""Builds the in-tree HIP library libads_mi355x.so for gfx950 with hipcc.

    python -m ads_mpi_amd.build          # compile (always recompiles)

The shared library lands in ads_mpi_amd/_lib/ so it travels with the source
tree; nothing is installed into site-packages.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT_DIR = os.path.join(PKG, "_lib")
LIB_NAME = "libads_mi355x.so"
LIB_PATH = os.path.join(OUT_DIR, LIB_NAME)

ARCH = os.environ.get("ADS_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "--offload-arch=" + ARCH,
    "-munsafe-fp-atomics",
    # packed FP32 VALU ops are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-variable",
    "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]

SOURCES = [
    ("setup.cpp", True),
    ("part_tables.cpp", False),
    ("assemble.hip", True),
    ("sweep.hip", True),
    ("part.hip", True),
    ("capi.cpp", True),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, extra_flags=None) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    objs = []
    cc = hipcc()
    flags = HIP_FLAGS + list(extra_flags or [])
    procs = []
    for src, is_hip in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OUT_DIR, src.replace(".", "_") + ".o")
        cmd = [cc] + flags + (["-x", "hip"] if is_hip else []) + ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    failed = []
    for src, pr in procs:
        out, err = pr.communicate()
        if pr.returncode != 0:
            failed.append("hipcc failed for %s:\n%s%s" % (src, out, err))
        elif verbose and err.strip():
            print(err, file=sys.stderr)
    if failed:
        raise RuntimeError("\n".join(failed))
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (res.stdout, res.stderr))
    os.replace(tmp, LIB_PATH)
    for o in objs:
        os.remove(o)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
