"""The fused-form host compiler (qvnt_amd/csrc/fusion.cpp) under address
and undefined-behaviour sanitizers: tests/cpp/fused_model.cpp models the
fused kernel on the CPU, runs the compiled programs of random circuits of
every gate kind (controls, daggers, dense phases, FP32-size tiles) and
compares them with the oracle's plain apply."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="needs gcc")
def test_fused_compiler_model_under_sanitizers(tmp_path):
    oracle_o = str(tmp_path / "oracle.o")
    subprocess.check_call(["gcc", "-std=gnu11", "-ffp-contract=off", "-fopenmp", *SAN, "-c",
                           os.path.join(ROOT, "oracle", "qvnt_oracle.c"), "-o", oracle_o])
    exe = str(tmp_path / "fused_model")
    subprocess.check_call(["g++", "-std=c++17", "-fopenmp", *SAN, "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "fused_model.cpp"),
                           os.path.join(ROOT, "qvnt_amd", "csrc", "fusion.cpp"),
                           os.path.join(ROOT, "qvnt_amd", "csrc", "ops.cpp"), oracle_o, "-o", exe])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="1")
    res = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failure(s)" in res.stdout
