"""The reference's own tests restated in C++ against include/qvnt.hpp."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "reference_tests.cpp")
LIBDIR = os.path.join(ROOT, "qvnt_amd", "lib")


@pytest.fixture(scope="module")
def binary(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "reference_tests")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"), SRC,
                           "-L", LIBDIR, "-lqvnt_hip", "-Wl,-rpath," + LIBDIR, "-o", out])
    return out


def test_operator_tests_host(binary):
    res = subprocess.run([binary, "host"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr


@pytest.mark.gpu
def test_register_tests_device(binary):
    res = subprocess.run([binary, "device"], capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
