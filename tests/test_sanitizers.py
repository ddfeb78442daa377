"""The host code (octree, keys, masks, immersed-boundary weights, kernel
tables: csrc/host/*.cpp and csrc/capi_host.cpp) and the FP64 oracle
(oracle/oracle.c) built with AddressSanitizer and UndefinedBehaviorSanitizer
and run through tests/sanitize_host.c: tree construction with refinement,
correction blocks and masks, export, the immersed-boundary stencil, and the
oracle's solves and transfers on the exported trees.  Any sanitizer report
aborts the run."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "iblgf-amr_amd", "csrc")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc/g++")
def test_host_code_and_oracle_are_sanitizer_clean(tmp_path):
    objs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp"))) + [os.path.join(CSRC, "capi_host.cpp")]:
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(["g++", "-std=c++17", "-c", src, "-o", obj] + SAN, check=True, capture_output=True)
        objs.append(obj)
    for src in (os.path.join(ROOT, "oracle", "oracle.c"), os.path.join(ROOT, "tests", "sanitize_host.c")):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(["gcc", "-std=c99", "-c", src, "-o", obj, "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "oracle")] + SAN, check=True, capture_output=True)
        objs.append(obj)
    exe = str(tmp_path / "sanitize_host")
    subprocess.run(["g++", "-o", exe] + objs + SAN + ["-lm"], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "sanitized host checks passed" in r.stdout
