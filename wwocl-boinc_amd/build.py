"""Build step of the package: libww.so (gfx950 kernels + C ABI + host checks) and the
drop-in host program bin/wwocl. Run from anywhere; outputs land next to this file.

    python wwocl-boinc_amd/build.py
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are kept off in every HIP compile.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-command-line-argument"]

LIB_SOURCES = ["ww_kernels.hip", "ww_batch.hip", "ww_presieve.hip", "ww_bucket.hip", "ww_api.cpp",
               "ww_server.cpp", "ww_cpu.cpp", "ww_selftest.cpp", "ww_peak.hip",
               "ww_noalloc.hip"]
HOST_SOURCES = ["wwocl_main.cpp", "wwocl_checks.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "argument unused" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_lib():
    out = os.path.join(HERE, "libww.so")
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = out + ".tmp"
    _run([HIPCC, "-x", "hip", *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_host():
    srcs = [os.path.join(HOST, f) for f in HOST_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    out = os.path.join(HERE, "bin", "wwocl")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o", out,
          *srcs, "-L", HERE, "-lww", "-Wl,-rpath,$ORIGIN/..", "-pthread"])
    return out


def build():
    lib = build_lib()
    host = build_host()
    return lib, host


if __name__ == "__main__":
    for p in build():
        if p:
            print("built", p)
    sys.exit(0)
