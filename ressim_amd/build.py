"""Build the HIP FIM library (libressim_fim.so) in-tree for gfx950.

Next to the library the build writes a stamp with the SHA-256 of every source it was
built from (the .hip files, every header under csrc/ and the public header). The loader
compares it with the tree, so a library built from other sources is never loaded as if
it were current: it is rebuilt, or, where no compiler is present, loading fails.
"""
from __future__ import annotations

import glob
import hashlib
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libressim_fim.so")
STAMP = LIB + ".sha256"
# translation units; every header under csrc/ and include/ is a dependency of all of them
SOURCES = ["rsf_lib.hip", "rsf_cnv_mb.hip"]


def dependencies() -> list[str]:
    deps = [os.path.join(CSRC, s) for s in SOURCES]
    deps += sorted(glob.glob(os.path.join(CSRC, "*.hpp")))
    deps += sorted(glob.glob(os.path.join(INCLUDE, "*.h")))
    return deps


def source_hash() -> str:
    h = hashlib.sha256()
    for path in dependencies():
        h.update(os.path.basename(path).encode())
        with open(path, "rb") as f:
            h.update(hashlib.sha256(f.read()).digest())
    return h.hexdigest()


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every HIP source into one shared library (always recompiles)."""
    digest = source_hash()
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           # keep packed FP32 instructions out of the generated code
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           # a declared but undefined function fails the build, not the first load
           "-Wl,--no-undefined",
           "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stderr[-4000:])
    if verbose and res.stderr:
        print(res.stderr[-2000:])
    os.replace(tmp, LIB)
    with open(STAMP, "w") as f:
        f.write(digest + "\n")
    return LIB


def needs_build() -> bool:
    """True when the library is absent or was built from sources other than the tree's."""
    if not os.path.exists(LIB) or not os.path.exists(STAMP):
        return True
    with open(STAMP) as f:
        return f.read().strip() != source_hash()


if __name__ == "__main__":
    print(build(verbose=True))
