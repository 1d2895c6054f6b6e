"""Build the native library and the C host in-tree.

    libstarrynight_amd.so  hipcc, gfx950: the HIP kernels + the C ABI
    sn_host                gcc: the C driver that reads starrynight.cfg

Run as ``python -m starrynight_amd.build`` (or through __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libstarrynight_amd.so")
HOST = os.path.join(PKG, "sn_host")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SN_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (no FP32 math on the hot path anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]

SOURCES = ["sn_api.hip", "sn_schedule.cpp", "sn_lattice_init.cpp", "sn_autotune.cpp", "sn_gather_kernels.hip"] + \
    [f"sn_field_kernels_rc{r}.hip" for r in range(1, 6)]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_lib(extra=(), jobs=None):
    """Compile every translation unit (in parallel) into _build/, link the .so."""
    import concurrent.futures as cf
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    cmds = []
    for s in SOURCES:
        o = os.path.join(bdir, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        cmds.append([HIPCC, *HIP_FLAGS, *extra, "-c", "-o", o, os.path.join(CSRC, s)])
    jobs = jobs or min(len(cmds), os.cpu_count() or 4, 16)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)


def build_host():
    hdir = os.path.join(CSRC, "host")
    srcs = [os.path.join(hdir, "sn_host.c"), os.path.join(hdir, "sn_config.c")]
    if not all(os.path.exists(s) for s in srcs):
        return
    _run(["gcc", "-O2", "-std=gnu11", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", hdir,
          "-o", HOST, *srcs, "-L", PKG, "-lstarrynight_amd", "-Wl,-rpath,$ORIGIN", "-lm"])


def build():
    build_lib()
    build_host()


if __name__ == "__main__":
    build()
    sys.exit(0)
