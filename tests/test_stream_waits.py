"""The streamed x kernel's hand-set wait, checked on its gfx950 assembly.

xphys_stream_kernel (channel_passive_amd/csrc/kernels.hip) copies the next
line into LDS with global_load_lds, then stores the current line, and at the
top of its loop waits with s_waitcnt vmcnt(STORES) rather than vmcnt(0).
Vector memory operations count down in issue order, so that wait covers the
copy only if at least STORES vector memory operations are issued after it on
every path to the wait.  This test compiles the kernel to gfx950 assembly
(device only, the library's own flags), builds the control-flow graph of
the kernel and checks that property over all paths, so that a compiler that
splits, merges or reorders those stores fails here instead of racing on the
GPU.  A wait that itself covers the copy (vmcnt(k) with k <= the operations
issued after it) ends a path.
"""
import heapq
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KERNELS = os.path.join(ROOT, "channel_passive_amd", "csrc", "kernels.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
STORES = 8  # the template argument of stream_sync in xphys_stream_kernel

TU = """#include "{kernels}"
template __global__ void pb::xphys_stream_kernel<512>(double2 *__restrict__, const double2 *__restrict__, int, int,
                                                      size_t, const double *__restrict__, double, double,
                                                      double *__restrict__);
"""

VMEM = re.compile(r"^\s*(global_|buffer_|flat_|scratch_)\w*")
COPY = re.compile(r"^\s*(global|buffer)_load_lds_")
WAIT = re.compile(r"^\s*s_waitcnt\b(.*)")
LABEL = re.compile(r"^(\.LBB\w+):")
BRANCH = re.compile(r"^\s*(s_branch|s_cbranch_\w+)\s+(\.LBB\w+)")


def vm_wait(operands: str):
    """vmcnt of an s_waitcnt line, or None when it does not wait on vmcnt."""
    if operands.strip() in ("", "0"):
        return 0 if operands.strip() == "0" else None
    m = re.search(r"vmcnt\((\d+)\)", operands)
    return int(m.group(1)) if m else None


def kernel_body(asm: str, name_part: str):
    lines = asm.splitlines()
    start = next(i for i, l in enumerate(lines) if re.match(r"^_Z\w*%s\w*:" % name_part, l))
    end = next(i for i in range(start, len(lines)) if "s_endpgm" in lines[i])
    return [l.split(";")[0].rstrip() for l in lines[start + 1:end + 1]]


def blocks_of(body):
    """Basic blocks: (label, [instructions], [successor labels])."""
    blocks, cur, insts = [], "entry", []
    for l in body:
        m = LABEL.match(l)
        if m:
            blocks.append([cur, insts])
            cur, insts = m.group(1), []
            continue
        if l.strip():
            insts.append(l)
    blocks.append([cur, insts])
    out = []
    for i, (lab, ins) in enumerate(blocks):
        succ = []
        last = ins[-1] if ins else ""
        m = BRANCH.match(last)
        nxt = blocks[i + 1][0] if i + 1 < len(blocks) else None
        if m:
            succ.append(m.group(2))
            if m.group(1) != "s_branch" and nxt:
                succ.append(nxt)
        elif "s_endpgm" not in last and nxt:
            succ.append(nxt)
        out.append((lab, ins, succ))
    return out


def ops_at_wait(blocks, k_wait):
    """Vector memory operations issued after a copy, for each shortest path
    from a copy to an s_waitcnt vmcnt(k_wait) that no earlier wait covers."""
    index = {lab: i for i, (lab, _, _) in enumerate(blocks)}
    # states: (operations issued since the copy, block, instruction)
    heap = [(0, bi, ii + 1) for bi, (_, ins, _) in enumerate(blocks) for ii, l in enumerate(ins) if COPY.match(l)]
    heapq.heapify(heap)
    seen, reached = set(), []
    while heap:
        n, bi, ii = heapq.heappop(heap)
        if (bi, ii) in seen:
            continue
        seen.add((bi, ii))
        _, ins, succ = blocks[bi]
        ended = False
        for l in ins[ii:]:
            w = WAIT.match(l)
            k = vm_wait(w.group(1)) if w else None
            if k is not None and (k == k_wait or n >= k):
                if k == k_wait:
                    reached.append(n)
                ended = True  # the wait covers the copy, or it is the one under test
                break
            if VMEM.match(l):
                n += 1
        if not ended:
            for s in succ:
                heapq.heappush(heap, (n, index[s], 0))
    return reached


@pytest.fixture(scope="module")
def stream_asm(tmp_path_factory):
    if not os.path.exists(HIPCC) and not shutil.which("hipcc"):
        pytest.skip("hipcc not present")
    d = tmp_path_factory.mktemp("stream_asm")
    src = d / "tu.hip"
    src.write_text(TU.format(kernels=KERNELS))
    out = d / "tu.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "--offload-device-only", "-S", str(src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


def test_stream_wait_covers_copy(stream_asm):
    body = kernel_body(stream_asm, "xphys_stream_kernel")
    waits = [l for l in body if WAIT.match(l) and vm_wait(WAIT.match(l).group(1)) == STORES]
    assert len(waits) == 1, f"expected one s_waitcnt vmcnt({STORES}) in the loop, found {waits}"
    blocks = blocks_of(body)
    assert any(COPY.match(l) for _, ins, _ in blocks for l in ins), "no LDS copies found"
    counts = ops_at_wait(blocks, STORES)
    assert counts, "no copy reaches the loop's wait"
    assert min(counts) >= STORES, f"a path reaches vmcnt({STORES}) with only {min(counts)} operations after a copy"


def test_stream_wait_is_tight(stream_asm):
    """Exactly STORES stores follow the copy on the shortest path: more would
    make the wait also hold stores of the previous line (slower, still correct)."""
    counts = ops_at_wait(blocks_of(kernel_body(stream_asm, "xphys_stream_kernel")), STORES)
    assert counts and min(counts) == STORES, f"operations between the copy and the wait: {sorted(set(counts))}"


def test_checker_sees_a_missing_store(stream_asm):
    """The checker itself: the same assembly with one store of the loop
    removed must show a path with STORES - 1 operations."""
    body = kernel_body(stream_asm, "xphys_stream_kernel")
    stores = [i for i, l in enumerate(body) if re.match(r"^\s*global_store_dwordx4", l)]
    mutant = body[:stores[-1]] + body[stores[-1] + 1:]
    assert min(ops_at_wait(blocks_of(mutant), STORES)) == STORES - 1
