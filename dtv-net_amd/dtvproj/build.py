"""Compile the gfx950 HIP library ``libdtvproj.so`` in-tree with hipcc.

Run as ``python -m dtvproj.build`` (with ``dtv-net_amd`` on ``sys.path``) or
call :func:`build`.  The library lands next to this file so that it travels
with the source tree.
"""
from __future__ import annotations

import glob
import hashlib
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
SRC_ROOT = os.path.dirname(PKG_DIR)
REPO_ROOT = os.path.dirname(SRC_ROOT)
CSRC = os.path.join(SRC_ROOT, "csrc")
INCLUDE = os.path.join(REPO_ROOT, "include")
LIB_NAME = "libdtvproj.so"
LIB_PATH = os.path.join(PKG_DIR, LIB_NAME)
ARCH = "gfx950"

SOURCES = ["dtv_project.hip", "dtv_brick.hip", "dtv_tv.hip", "dtv_fbp.hip", "dtv_shrink.hip"]

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 forms stay off: compiler-generated only, never relied upon
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc_path() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (needed to build the gfx950 library)")


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def digest_inputs():
    """Every file the library is compiled from: the HIP sources, the private
    headers and the public C header."""
    files = glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.h"))
    files += glob.glob(os.path.join(INCLUDE, "*.h"))
    return sorted(files)


def source_digest() -> str | None:
    """First 16 hex digits of the SHA-256 over the library's inputs (name and
    contents of each file), or None when the sources are not present.  The
    build compiles it into ``dtv_version()``, and the loader refuses a library
    whose digest differs from that of the sources beside it."""
    files = digest_inputs()
    if not files:
        return None
    h = hashlib.sha256()
    for f in files:
        h.update(os.path.basename(f).encode() + b"\0")
        with open(f, "rb") as fh:
            h.update(fh.read())
        h.update(b"\0")
    return h.hexdigest()[:16]


def build(verbose: bool = False) -> str:
    """Compile every HIP source into ``libdtvproj.so``; returns its path."""
    out_tmp = LIB_PATH + ".tmp"
    cmd = [hipcc_path(), *HIPCC_FLAGS, f"-DDTV_SOURCE_DIGEST={source_digest()}", f"-I{CSRC}",
           f"-I{INCLUDE}", *sources(), "-o", out_tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{proc.stdout}")
    if verbose and proc.stdout.strip():
        print(proc.stdout)
    os.replace(out_tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
