"""The transposed wall-normal kernels stay within the register file.

adj_chain_kernel<4> and adj_hvhg_part_kernel<4> batch their rows
(ADJ_BATCH in solver_adj.h) to overlap memory latency.  A batch of 16 rows
pushed both to 256 VGPRs plus AGPRs and the chain kernel to 68 bytes per
lane of scratch, and that build faulted on the GPU; the batch of 8 (211 and
205 VGPRs, no scratch) is the one that runs.  This CPU test reads the
compiler's resource report of the built library and fails if either
spills to scratch or reaches the 256-VGPR limit again (from the report the
build writes).
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REPORT = os.path.join(ROOT, "channel_passive_amd", "kernel_resources.txt")


@pytest.fixture(scope="module")
def remarks():
    """The compiler's resource report of the built library (written by the
    build next to libpoongback_hip.so)."""
    if not os.path.exists(REPORT):
        pytest.skip("library not built here")
    return open(REPORT).read()


def usage(remarks, name):
    lines = remarks.splitlines()
    i = next(k for k, l in enumerate(lines) if "Function Name:" in l and name in l)
    out = {}
    for l in lines[i + 1:i + 16]:
        if "Function Name:" in l:
            break
        m = re.search(r"remark:\s+(VGPRs|AGPRs|ScratchSize) ?(?:\[bytes/lane\])?: (\d+)", l)
        if m:
            out[m.group(1)] = int(m.group(2))
    return out


@pytest.mark.parametrize("name", ["adj_chain_kernel", "adj_hvhg_part_kernel"])
def test_adjoint_kernels_do_not_spill(remarks, name):
    u = usage(remarks, name)
    assert u["ScratchSize"] == 0 and u["AGPRs"] == 0 and u["VGPRs"] < 256, u
