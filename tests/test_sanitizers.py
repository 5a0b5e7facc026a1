"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host shell (XML /
OBJ loading, procedural scenes, parallel BVH build, flattening, EXR reading)
and of the oracle (scene build and render), on the CPU.

The host sources and the oracle are compiled with g++ (they contain no device
code) together with tests/sanitize_host.cpp into one executable, which loads
the reference's preview scene file, generates each workload scene at a small
size, and reads an EXR file written by the test's own writer.  A sanitizer
report aborts the run; the test requires a clean exit and no report."""
from __future__ import annotations

import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from imgio import write_exr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = ["hmath", "bvh_build", "scene_host", "generators", "xml_loader", "exr_read"]
ORACLE = ["gnd_ref_bvh", "gnd_ref_render", "gnd_ref_capi", "gnd_ref_denoise", "gnd_ref_motion"]
FLAGS = ["-O1", "-g", "-std=c++17", "-pthread", "-fno-omit-frame-pointer", "-ffp-contract=off",
         "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


@pytest.fixture(scope="module")
def sanitized_exe(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("asan")
    srcs = [os.path.join(ROOT, "gianduia_amd", "csrc", "host", f + ".cpp") for f in HOST]
    srcs += [os.path.join(ROOT, "oracle", f + ".cpp") for f in ORACLE]
    srcs.append(os.path.join(ROOT, "tests", "sanitize_host.cpp"))

    def compile_one(src):
        obj = os.path.join(out, os.path.basename(src) + ".o")
        r = subprocess.run([cxx, *FLAGS, "-c", src, "-o", obj], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{src}: {r.stderr[-2000:]}")
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, srcs))
    exe = os.path.join(out, "sanitize_host")
    r = subprocess.run([cxx, *FLAGS, *objs, "-lz", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe, out


def test_host_shell_and_oracle_under_asan_ubsan(sanitized_exe):
    exe, out = sanitized_exe
    exr = os.path.join(out, "sky.exr")
    rng = np.random.default_rng(3)
    img = rng.random((12, 20, 3)).astype(np.float32)
    write_exr(exr, {"R": img[..., 0], "G": img[..., 1], "B": img[..., 2]}, ptype=1, compression=3)
    xml = os.path.join(ROOT, "tests", "golden", "reference_cbox", "cbox.xml")
    # the environment may preload a library ahead of the ASan runtime
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, xml, exr], capture_output=True, text=True, env=env, timeout=600)
    report = r.stdout[-3000:] + r.stderr[-6000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert r.stdout.count("film weight") == 7, report
    assert "exr: 20x12" in r.stdout, report
