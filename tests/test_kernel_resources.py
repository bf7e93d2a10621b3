"""Register allocation of the fused kernels, from the compiler's own
resource report: the 16 amplitudes each thread holds must stay in VGPRs
(no scratch memory), for both precisions and both launch forms."""
import os
import re
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "qvnt_amd"))
import build  # noqa: E402  (qvnt_amd/build.py: the library's own compile flags)

sys.path.pop(0)


@pytest.mark.skipif(not os.path.exists(build.HIPCC), reason="needs hipcc")
def test_fused_kernels_do_not_spill(tmp_path):
    cmd = build.compile_command("fused.hip", str(tmp_path / "fused.o")) + ["-Rpass-analysis=kernel-resource-usage"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr
    kernels = {}
    name = None
    for line in res.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            kernels[name] = int(m.group(1))
    fused = {k: v for k, v in kernels.items() if "k_fused" in k or "k_batch" in k}
    assert len(fused) == 4, kernels
    assert all(v == 0 for v in fused.values()), fused
