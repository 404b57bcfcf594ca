"""Build-level checks of the solver kernel (CPU only: hipcc cross-compiles for gfx950).

The persistent kernel must not use scratch (private) memory: every spill or stack object there is an
uncoalesced round trip to memory on the hot path, and one slipped in unnoticed twice while the kernel
grew (a struct passed by reference to an out-of-line function, a local whose address was taken)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_solver_kernel_uses_no_scratch(tmp_path):
    src = os.path.join(ROOT, "adds_amd", "csrc", "sssp_kernels.hip")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-c", src,
           "-o", str(tmp_path / "k.o"), "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    # remarks come per kernel: "Function Name: <mangled>" followed by its resource lines
    current, scratch = None, {}
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            current = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and current:
            scratch[current] = int(m.group(1))
    solvers = {k: v for k, v in scratch.items() if "k_sssp" in k}
    assert len(solvers) == 8, scratch  # int/float x single/replicated x 4/8 edge windows
    assert all(v == 0 for v in solvers.values()), solvers
