"""Builds libqvnt_hip.so in-tree with hipcc for gfx950.

Usage: python qvnt_amd/build.py  (or __graft_entry__.build()); it does not
import the package, which loads the library it builds.
Objects go to qvnt_amd/_build/, the library to qvnt_amd/lib/.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ_DIR = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libqvnt_hip.so")

SOURCES = ["ops.cpp", "fusion.cpp", "plain.hip", "measure.hip", "fused.hip", "reg.hip", "serve.hip", "batch.hip", "algebra.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: the plain kernels restate the reference's arithmetic
# operation by operation; packed FP32 instructions are disabled for all
# device code.
COMMON = ["-std=c++17", "-O3", "-fPIC", "-Wall", "-Wno-unused-result",
          "--offload-arch=" + ARCH, "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


# fused.hip: the op dispatch switches over a register bit, and LLVM's
# switch-tail sinking merges the cases into one block that indexes the
# 16-amplitude register array dynamically, which sends the array to
# scratch memory; without the sinking every case keeps constant indices.
EXTRA = {"fused.hip": ["-mllvm", "-simplifycfg-sink-common=false"]}


def compile_command(src, obj):
    """The hipcc command line that builds one source file."""
    return [HIPCC] + COMMON + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = compile_command(src, obj)
    if src.endswith(".cpp"):
        cmd = [HIPCC, "-x", "c++"] + [f for f in COMMON if not f.startswith("--offload-arch")
                                      and f not in ("-Xclang", "-target-feature", "-packed-fp32-ops")] + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), res.stdout, res.stderr))
    return obj, res.stderr


def build(verbose=False, jobs=None):
    """Compile every source and link the shared library (always rebuilds)."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for obj, err in results:
            if err.strip():
                sys.stderr.write(err)
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + [o for o, _ in results]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), res.stderr))
    shutil.move(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
