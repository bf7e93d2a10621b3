"""The C ABI library loads and exports every entry point include/qvnt.h
declares; host-only calls work without a GPU."""
import ctypes
import os
import re

import pytest

import qvnt_amd
from qvnt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "qvnt.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qvnt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "qvnt_reg_apply" in names and "qvnt_reg_measure_mask" in names
    assert len(names) >= 50


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_gate_layout():
    assert b"gfx950" in _lib.lib.qvnt_version()
    assert ctypes.sizeof(qvnt_amd.Gate) == 312


def test_register_needs_a_gfx950_device():
    """Without a visible gfx950 device the register refuses (no CPU path)."""
    import subprocess
    import sys
    code = ("import qvnt_amd, sys\n"
            "try:\n    qvnt_amd.QReg(4)\nexcept qvnt_amd.QvntError as e:\n    sys.exit(0 if e.code == -7 else 1)\n"
            "sys.exit(2)\n")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    res = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, timeout=120)
    assert res.returncode == 0, res.stderr.decode()[-2000:]
