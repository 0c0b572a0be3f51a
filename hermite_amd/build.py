"""Build recipe for the in-tree native libraries.

``libhermite_hip.so``   gfx950 kernels + the C-ABI of include/hermite_hip.h
``libhermite_host.so``  host-side inputs (Plummer/Salpeter generator) and the
                        block-step driver that calls the C-ABI
``hermite_run``         command-line driver reading the reference's parameter
                        tag files (built when its source is present)
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")

HIP_LIB = os.path.join(PKG, "libhermite_hip.so")
HOST_LIB = os.path.join(PKG, "libhermite_host.so")
RUNNER = os.path.join(PKG, "hermite_run")

ARCH = "gfx950"
# -ffp-contract=off: every FMA in the kernels is explicit, so the EXACT
# variant keeps the reference's rounding.  Packed FP32 is disabled on purpose.
HIPFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
]
HOSTFLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-Wall"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; the ROCm toolchain is required to build hermite_amd")


def _run(cmd, quiet=True):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _atomic(out: str, cmd_for_tmp):
    tmp = out + ".tmp"
    _run(cmd_for_tmp(tmp))
    os.replace(tmp, out)


def build(verbose: bool = False) -> None:
    """Compile every native artefact of the package in-tree."""
    hip_srcs = [os.path.join(CSRC, "hh_gravity.hip")]
    _atomic(HIP_LIB, lambda o: [_hipcc()] + HIPFLAGS + ["-shared", "-o", o] + hip_srcs)
    host_srcs = [os.path.join(CSRC, f) for f in ("hh_host.cpp", "hh_driver.cpp") if os.path.exists(os.path.join(CSRC, f))]
    _atomic(HOST_LIB, lambda o: ["g++"] + HOSTFLAGS + ["-I" + INCLUDE, "-shared", "-o", o] + host_srcs
            + ["-L" + PKG, "-lhermite_hip", "-Wl,-rpath,$ORIGIN"])
    main = os.path.join(CSRC, "hh_main.cpp")
    if os.path.exists(main):
        _atomic(RUNNER, lambda o: ["g++"] + HOSTFLAGS + ["-I" + INCLUDE, "-o", o, main]
                + ["-L" + PKG, "-lhermite_host", "-lhermite_hip", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", HIP_LIB, HOST_LIB)


if __name__ == "__main__":
    build(verbose=True)
