"""Control flow of the persistent batched kernel as gfx950 code (CPU only:
hipcc cross-compiles the device code to assembly).

batched_kernel loops over work items: thread 0 takes the next item from a
global counter and leaves it in LDS, a barrier publishes it, and the item ends
with the block reduction, whose last step runs on thread 0 alone.  When thread
0's reduction tail and its next fetch met across the loop's back edge, the
compiler threaded thread 0 from one to the other and built a loop nest in which
thread 0 leaves an inner loop that still holds the loop's barriers: the other
lanes of wave 0 and the other waves passed those barriers without it, re-read
the old item and never ended (two forms with the tables in global memory
stalled this way).  The kernel now ends every item with a barrier; this test
checks in the generated code that every barrier of every instance sits in the
item loop itself (loop depth 1), as the compiler's loop annotations give it.
"""
import os
import re
import subprocess

import pytest

from twisty_amd import build as B


def _hipcc():
    try:
        return B.hipcc()
    except RuntimeError:
        return None


HIPCC = _hipcc()


def barrier_depths(asm):
    """{kernel symbol: [loop depth of each s_barrier]} for the batched_kernel
    instances of an assembly listing."""
    out, cur, depth = {}, None, 0
    for line in asm.splitlines():
        m = re.match(r"^(_ZN10twisty_hip14batched_kernel\w+):", line)
        if m:
            cur, depth = m.group(1), 0
            out[cur] = []
            continue
        if cur is None:
            continue
        if "Loop Header: Depth=" in line:
            depth = int(line.split("Depth=")[1].split()[0])
        elif re.match(r"^\.LBB\d+_\d+:", line) or line.startswith("; %bb"):
            d = re.search(r"Depth=(\d+)", line)
            depth = int(d.group(1)) if d else 0
        if "s_barrier" in line:
            out[cur].append(depth)
        if "s_endpgm" in line:
            cur = None
    return out


@pytest.mark.skipif(HIPCC is None, reason="hipcc not available")
def test_batched_barriers_stay_in_the_item_loop(tmp_path):
    asm = tmp_path / "batched.s"
    cmd = [HIPCC, *B.HIPCC_FLAGS, "--cuda-device-only", "-S", os.path.join(B.CSRC, "batched.hip"), "-o", str(asm)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    depths = barrier_depths(asm.read_text())
    print({k[34:48]: v for k, v in depths.items()}, flush=True)
    # every placement (table in LDS, lean, table in global memory) is compiled
    assert len(depths) >= 5 and {re.search(r"ELi(\d)EEEv", k).group(1) for k in depths} == {"0", "1", "2"}
    for sym, d in depths.items():
        assert len(d) >= 4 and set(d) == {1}, (sym, d)


def test_barrier_depths_reads_nested_loops():
    """The parser on a listing shaped like the stalled form: the barriers in an
    inner loop (depth 2) are reported as such."""
    asm = """_ZN10twisty_hip14batched_kernelILi256ELi1EEEvNS_11BatchParamsE:
.LBB4_4:                                ; =>This Loop Header: Depth=1
\ts_cbranch_execz .LBB4_8
.LBB4_11:                               ;   Parent Loop BB4_4 Depth=1
                                        ; =>  This Loop Header: Depth=2
\ts_barrier
; %bb.12:                               ;   in Loop: Header=BB4_11 Depth=2
\ts_barrier
.LBB4_611:                              ;   in Loop: Header=BB4_4 Depth=1
\ts_barrier
.LBB4_700:
\ts_endpgm
"""
    assert barrier_depths(asm) == {"_ZN10twisty_hip14batched_kernelILi256ELi1EEEvNS_11BatchParamsE": [2, 2, 1]}
