"""Build property of the HIP sources: no kernel uses scratch (private) memory.

Register spills and locally copied by-value arrays put per-lane data in scratch
memory, which the runtime must back per dispatch; the engine's kernels are
written to stay in registers, LDS and plain device buffers.  The compiler's
resource report for gfx950 (-Rpass-analysis=kernel-resource-usage) is read for
every kernel of every source file of libesp.so.
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "esys-particle_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["esp_kernels.hip", "esp_engine.hip", "esp_diff.hip"]


def resource_report(src, tmpdir):
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only",
           "-c", os.path.join(CSRC, src), "-o", os.path.join(tmpdir, src + ".o"),
           "-Rpass-analysis=kernel-resource-usage"]
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    kernels = {}
    name = None
    for line in r.stderr.splitlines():
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            continue
        m = re.search(r"remark:\s+ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            kernels[name] = int(m.group(1))
    return kernels


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not found")
def test_no_kernel_uses_scratch(tmp_path):
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        reports = list(ex.map(lambda s: resource_report(s, str(tmp_path)), SOURCES))
    kernels = {k: v for rep in reports for k, v in rep.items()}
    # the per-step kernels of both engines are among them
    assert any("k_forces" in k for k in kernels) and any("k_bond_fwd" in k for k in kernels)
    assert any("k_gather_integrate" in k for k in kernels)
    bad = {k: v for k, v in kernels.items() if v != 0}
    assert not bad, "kernels using scratch memory: %s" % bad
