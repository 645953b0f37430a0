"""Build recipe of the native libraries (run by __graft_entry__.build()).

Everything is compiled in-tree so the shared objects travel with the tree:
  mgbrain_amd/lib/libmgbrain_hip.so   HIP kernels + C ABI (hipcc, gfx950)
  mgbrain_amd/lib/libmgbrain.so       C++ facade mirroring MGBrain's API (g++)
  mgbrain_amd/lib/mgb_*               C++ example / driver programs
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")

HIP_SOURCES = ["mgb_kernels.hip", "mgb_build.hip", "mgb_api.hip"]
HOST_SOURCES = ["host/model.cpp", "host/netgen.cpp", "host/mgsim.cpp"]
PROGRAMS = {"mgb_run_model": "host/run_model.cpp"}

ARCH = os.environ.get("MGB_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: every fused multiply-add in the kernels is explicit.
# -packed-fp32-ops: no packed FP32 instructions are generated.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
             "-Wno-unused-function", f"-I{INC}", f"-I{CSRC}"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-sign-compare",
             f"-I{INC}", f"-I{CSRC}"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip(jobs=8):
    os.makedirs(os.path.join(LIB, "obj"), exist_ok=True)

    def one(src):
        obj = os.path.join(LIB, "obj", os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    out = os.path.join(LIB, "libmgbrain_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_host():
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if not srcs:
        return None
    out = os.path.join(LIB, "libmgbrain.so")
    tmp = out + ".tmp"
    _run(["g++", *CXX_FLAGS, "-shared", "-o", tmp, *srcs, f"-L{LIB}", "-lmgbrain_hip",
          f"-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    for name, src in PROGRAMS.items():
        p = os.path.join(CSRC, src)
        if not os.path.exists(p):
            continue
        _run(["g++", *CXX_FLAGS, "-o", os.path.join(LIB, name), p, f"-L{LIB}", "-lmgbrain",
              "-lmgbrain_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/src/model"):
        _run(["bash", os.path.join(ROOT, "oracle", "build_ref.sh")])


def build_all():
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        f_or = ex.submit(build_oracle)
        build_hip()
        f_or.result()
    build_host()
    shutil.rmtree(os.path.join(LIB, "obj"), ignore_errors=True)


if __name__ == "__main__":
    build_all()
    print("built", LIB)
