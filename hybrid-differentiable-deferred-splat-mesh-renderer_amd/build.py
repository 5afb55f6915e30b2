"""Builds the HIP compositor library (lib/libhdsr.so) and the Node N-API addon for gfx950.

Every .hip file under csrc/ is compiled with hipcc and every .cpp file (host-only code: the asset
decoders) with g++, then the objects are linked into one shared library.  Floating-point
contraction is off (the kernels must round exactly as the host-side restatement does) and packed
FP32 instructions are disabled.

    python build.py            # library + addon
    python build.py --lib      # library only
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libhdsr.so")
NODE_DIR = os.path.join(HERE, "node")
ADDON = os.path.join(NODE_DIR, "hdsr_napi.node")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")

ARCH = os.environ.get("HDSR_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{CSRC}",
    f"-I{INCLUDE}",
]


CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", f"-I{CSRC}",
             f"-I{INCLUDE}"]


def _sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    if src.endswith(".cpp"):
        cmd = ["g++", *CXX_FLAGS, "-c", src, "-o", obj]
    else:
        cmd = [HIPCC, *HIP_FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_lib(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_addon() -> str | None:
    """Compiles the N-API addon against Node's headers; returns None when Node headers are absent."""
    node_inc = "/usr/include/node"
    src = os.path.join(NODE_DIR, "hdsr_napi.c")
    if not (os.path.isdir(node_inc) and os.path.exists(src)):
        return None
    cmd = ["gcc", "-O2", "-shared", "-fPIC", "-std=c11", f"-I{node_inc}", f"-I{INCLUDE}", "-o", ADDON, src,
           f"-L{LIB_DIR}", "-lhdsr", "-Wl,-rpath,$ORIGIN/../lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"addon build failed:\n{r.stdout}\n{r.stderr}")
    return ADDON


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build_lib(a.j))
    if not a.lib:
        print(build_addon())
    return 0


if __name__ == "__main__":
    sys.exit(main())
