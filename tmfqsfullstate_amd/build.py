"""Builds the native libraries in-tree (gfx950):

  lib/libtmfqs_hip.so  -- HIP kernels + C-ABI (include/tmfqs_hip.h)
  lib/libtmfqs.so      -- the C++ API (RegisterConfig / QuantumRegister /
                          circuit::Circuit / sim::Simulator with the Hip
                          storage strategy) over libtmfqs_hip.so
  bin/*                -- C++ test drivers mirroring the reference's ctest
                          executables

Usage: python -m tmfqsfullstate_amd.build  (or build() from __graft_entry__).
Every call recompiles.
"""

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
CLI = os.path.join(PKG, "cli")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Every HIP translation unit: gfx950 only, no packed-FP32 instruction
# selection, and no floating-point contraction (the per-operation kernels
# reproduce the reference's FP64 arithmetic bit for bit).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        print(out)
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd[:3])} ...")


def build_hip():
    """Compiles every HIP translation unit (in parallel) and links the C-ABI
    library."""
    os.makedirs(LIB, exist_ok=True)
    obj_dir = os.path.join(PKG, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    out = os.path.join(LIB, "libtmfqs_hip.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith((".hip", ".cpp"))]
    objs = [os.path.join(obj_dir, os.path.basename(s) + ".o") for s in srcs]
    jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", f"-I{INCLUDE}", f"-I{CSRC}", so[0], "-o", so[1]]),
                    zip(srcs, objs)))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_host():
    """C++ API library and the C++ test drivers (built with hipcc's host
    compiler; no device code)."""
    if not os.path.isdir(HOST):
        return None
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(LIB, "libtmfqs.so")
    tmp = out + ".tmp"
    srcs = []
    for dp, _, fs in os.walk(HOST):
        srcs += [os.path.join(dp, f) for f in fs if f.endswith(".cpp") and "tests" not in dp]
    cxx = ["/opt/rocm/llvm/bin/clang++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
           f"-I{INCLUDE}", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]
    _run([*cxx, "-shared", *sorted(srcs), "-o", tmp, f"-L{LIB}", "-ltmfqs_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    # The C++ test drivers and the command-line programs, one executable each.
    for d in (os.path.join(HOST, "tests"), CLI):
        if not os.path.isdir(d):
            continue
        for f in sorted(os.listdir(d)):
            if f.endswith(".cpp"):
                exe = os.path.join(BIN, f[:-4])
                _run([*cxx, os.path.join(d, f), "-o", exe + ".tmp", f"-L{LIB}", "-ltmfqs", "-ltmfqs_hip",
                      "-Wl,-rpath,$ORIGIN/../lib"])
                os.replace(exe + ".tmp", exe)
    return out


def build_oracle():
    """Compiles the test oracle (and, where the reference sources exist, the
    reference build) -- checker infrastructure, not the product."""
    make = shutil.which("make")
    if make is None:
        return
    _run([make, "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all():
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
