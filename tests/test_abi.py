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


def _c_layouts(tmp_path, structs):
    """sizeof and offsetof of every field of the header's structs, from the
    host C compiler."""
    lines = ['#include <stdio.h>', '#include "qvnt.h"', "int main(void) {"]
    for cname, fields in structs:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in fields:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


@pytest.mark.skipif(not os.path.exists("/usr/bin/gcc"), reason="needs gcc")
def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror of a header struct has the C layout: the same
    field names, offsets and size."""
    from qvnt_amd import batch, serve
    mirrors = [("qvnt_gate", _lib.Gate), ("qvnt_kstats", _lib.KStats), ("qvnt_step", _lib.Step),
               ("qvnt_batch_info", batch.BatchInfo), ("qvnt_request", serve.Request),
               ("qvnt_result", serve.Result)]
    structs = [(c, [f[0] for f in py._fields_]) for c, py in mirrors]
    c = _c_layouts(tmp_path, structs)
    for cname, py in mirrors:
        assert ctypes.sizeof(py) == c[cname], cname
        for f in py._fields_:
            assert getattr(py, f[0]).offset == c["%s.%s" % (cname, f[0])], (cname, f[0])


def test_cache_trim_without_a_device():
    """Emptying the register cache needs no device and leaves none behind."""
    qvnt_amd.cache_trim()
    qvnt_amd.cache_trim()
