"""In-tree build of the native libraries (HIP kernels for gfx950 + host C++), stamped by source hash.

Mirrors the reference's rebuild-on-source-change loader (vulcan/gpu/_lib.py:40-60,
scripts/m2/build_kernels.sh:12-16) with hipcc instead of nvcc:
  * libvulcan_hip.so  <- csrc/*.hip  (hipcc --offload-arch=gfx950, no FMA contraction, no packed-FP32 ops)
  * libvulcan_l1.so   <- csrc/l1_expand.cpp (g++ -O3 -fopenmp; peptidoform expansion, xxHash64)
"""
from __future__ import annotations

import glob
import hashlib
import os
import shutil
import subprocess

PKG = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
ROOT = os.path.abspath(os.path.join(PKG, ".."))
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(PKG, "lib")

HIP_SO = os.path.join(LIBDIR, "libvulcan_hip.so")
L1_SO = os.path.join(LIBDIR, "libvulcan_l1.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", INCLUDE, "-I", CSRC]


def _sources(kind: str) -> list[str]:
    if kind == "hip":
        return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.h")) +
                      glob.glob(os.path.join(INCLUDE, "*.h")))
    return [os.path.join(CSRC, "l1_expand.cpp")]


def _hash(files: list[str], extra: str) -> str:
    h = hashlib.sha256(extra.encode())
    for f in files:
        with open(f, "rb") as fh:
            h.update(os.path.basename(f).encode())
            h.update(fh.read())
    return h.hexdigest()


def _stamp(so: str) -> str:
    return so + ".sha256"


def is_fresh(kind: str) -> bool:
    so = HIP_SO if kind == "hip" else L1_SO
    flags = " ".join(HIP_FLAGS) if kind == "hip" else "g++"
    if not os.path.exists(so) or not os.path.exists(_stamp(so)):
        return False
    with open(_stamp(so)) as fh:
        return fh.read().strip() == _hash(_sources(kind), flags)


def build_l1(force: bool = False) -> str:
    if not force and is_fresh("l1"):
        return L1_SO
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = L1_SO + ".partial"
    subprocess.run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-o", tmp,
                    os.path.join(CSRC, "l1_expand.cpp")], check=True)
    os.replace(tmp, L1_SO)
    with open(_stamp(L1_SO), "w") as fh:
        fh.write(_hash(_sources("l1"), "g++"))
    return L1_SO


def build_hip(force: bool = False, verbose: bool = False) -> str:
    if not force and is_fresh("hip"):
        return HIP_SO
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found: cannot build libvulcan_hip.so")
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    tmp = HIP_SO + ".partial"
    cmd = [HIPCC] + HIP_FLAGS + ["-o", tmp] + srcs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout[-4000:] + r.stderr[-8000:])
    if verbose and r.stderr:
        print(r.stderr[-4000:])
    os.replace(tmp, HIP_SO)
    with open(_stamp(HIP_SO), "w") as fh:
        fh.write(_hash(_sources("hip"), " ".join(HIP_FLAGS)))
    return HIP_SO


def build_all(force: bool = True, verbose: bool = False) -> None:
    build_l1(force=force)
    build_hip(force=force, verbose=verbose)
