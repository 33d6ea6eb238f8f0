"""The library's host-side C++ (2-opt, orbit, piece and box search drivers,
move and instance generators) built with g++ under AddressSanitizer and
UndefinedBehaviorSanitizer, with the oracle as the caller-supplied exact
evaluator and stand-ins for the GPU entry points (tests/sanitize/).  CPU
only: no GPU code is built or run."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "graphgolf-cuda_amd", "csrc")
SAN = os.path.join(REPO, "tests", "sanitize")
FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="needs gcc/g++")
def test_host_code_under_sanitizers(tmp_path):
    oracle_o = tmp_path / "gg_oracle.o"
    subprocess.run(["gcc", "-std=gnu11", "-fopenmp", *FLAGS, "-c",
                    os.path.join(REPO, "oracle", "gg_oracle.c"), "-o", str(oracle_o)],
                   check=True, capture_output=True)
    exe = tmp_path / "host_driver"
    srcs = [os.path.join(CSRC, f) for f in ("gg_search.cpp", "gg_orbit.cpp", "gg_gen.cpp",
                                            "gg_inittemp.cpp")]
    srcs += [os.path.join(SAN, "gpu_entry_stubs.cpp"), os.path.join(SAN, "host_driver.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-fopenmp", *FLAGS,
                        "-I" + os.path.join(REPO, "include"), "-I" + CSRC, *srcs,
                        str(oracle_o), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "OK" in r.stdout and "runtime error" not in out and "ERROR: " not in out, out[-4000:]
