"""Builds libgnd_hip.so (HIP kernels for gfx950 + the host scene shell).

Usage: python -m gianduia_amd.build [--jobs N]

The device translation unit is compiled with contraction off, correctly
rounded fp32 division/sqrt and fp32 denormals kept, so BVH traversal
reproduces the CPU reference bit for bit.  Packed-fp32 instruction selection
is disabled for device code.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgnd_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GND_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
          "-I", os.path.join(ROOT, "include")]
DEVICE = [f"--offload-arch={ARCH}", "-fno-gpu-flush-denormals-to-zero",
          "-fhip-fp32-correctly-rounded-divide-sqrt",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]

HOST_SOURCES = ["host/hmath.cpp", "host/bvh_build.cpp", "host/scene_host.cpp", "host/generators.cpp",
                "host/xml_loader.cpp", "host/integrator.cpp", "host/exr_read.cpp", "host/host_capi.cpp"]
DEVICE_SOURCES = ["device/kernels.hip", "device/denoise.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    noisy = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def build(jobs: int | None = None, verbose: bool = False) -> str:
    """Compiles every source and links libgnd_hip.so; returns its path."""
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    os.makedirs(BUILD, exist_ok=True)
    tasks = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        tasks.append([HIPCC, "-x", "c++", *COMMON, "-pthread", "-c", os.path.join(CSRC, src), "-o", obj])
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        tasks.append([HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj])
    jobs = jobs or min(8, len(tasks))
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, tasks))
    objs = [t[-1] for t in tasks]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-pthread", *objs, "-lz", "-o", tmp])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs, verbose=True)


if __name__ == "__main__":
    main()
