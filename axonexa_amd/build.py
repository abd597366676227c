"""Build the HIP library (gfx950) in-tree.

    python -m axonexa_amd.build

Every HIP source is compiled with -ffp-contract=off (fused multiply-adds in
the walk are explicit fmaf) and with packed FP32 disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AXX_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

SOURCES = (["capi.hip", "batch.hip", "mlp.hip", "mag.hip", "mag_rotate.hip", "mag_rk4.hip", "mag_dopri.hip",
            "mag_batch.hip", "mag_batch_rotate.hip", "mag_batch_rk4.hip", "mag_batch_dopri.hip"]
           + [f"walk_p{p}.hip" for p in (1, 2, 4, 8, 16, 32, 64)])
LIBNAME = "libaxonexa.so"


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l
                    and "(ignoring feature)" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_lib(verbose: bool = False) -> str:
    """Compile every .hip source to an object and link libaxonexa.so."""
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{out}")
    lib = os.path.join(HERE, LIBNAME)
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, lib)
    return lib


DRIVERS = ["cos_onecyl", "gammadist", "wccompare", "maglf", "cos_onesphere", "hex_phase", "plane_phase"]


def build_drivers() -> list:
    """Compile the ported example drivers (C++ over include/axonexa_compat.hpp)."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    out = []
    for name in DRIVERS:
        exe = os.path.join(bindir, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
              os.path.join(HERE, "drivers", name + ".cpp"), "-o", exe,
              "-L", HERE, "-laxonexa", "-Wl,-rpath,$ORIGIN/.."])
        out.append(exe)
    return out


def build_oracle() -> str:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def main(argv=None):
    t = time.time()
    path = build_lib()
    build_drivers()
    print(f"built {path} and drivers in {time.time() - t:.1f}s")


if __name__ == "__main__":
    main(sys.argv[1:])
