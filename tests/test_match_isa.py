"""Loop shape of the matcher's persistent kernel in the gfx950 ISA (CPU test:
hipcc cross-compiles, nothing runs on a GPU).

Each wave of match_kernel loops: lane 0 takes a task index from the queue
with a returning atomic, the index is broadcast to the wave (ds_bpermute
from lane 0), and the wave exits when the queue is drained.  Both must sit
in the same outermost loop.  When the compiler instead nests a second loop
around the broadcast whose exit lane 0 takes alone (it did so behind
lane-0-only branches at the end of the task), the other lanes keep reading
task indices from the inactive lane 0 and never leave: a device hang.  This
test compiles csrc/match.hip and checks every instantiation (count, plan
trial, root log, budgeted count; binary and merge probes)."""
import os
import re
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "gpm-artifact_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def loop_shapes(asm_lines):
    """{kernel: (depth of the queue fetch, depth of its broadcast)}."""
    starts = [i for i, l in enumerate(asm_lines)
              if re.match(r"^_ZN12_GLOBAL__N_112match_kernel.*:", l)]
    out = {}
    for s in starts:
        name = asm_lines[s].split(":")[0]
        depth, fetch, bcast = 0, None, None
        for line in asm_lines[s + 1:]:
            if line.startswith(".Lfunc_end"):
                break
            m = re.search(r"=>.*Loop Header: Depth=(\d+)", line)
            if m:
                depth = int(m.group(1))
            if fetch is None and re.search(
                    r"global_atomic_add v\d+, v\d+, v\d+, s\[\d+:\d+\] sc0", line):
                fetch = depth
            if bcast is None and "ds_bpermute" in line:
                bcast = depth
            if fetch is not None and bcast is not None:
                break
        out[name] = (fetch, bcast)
    return out


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_match_kernel_task_loop_shape(tmp_path):
    asm = tmp_path / "match.s"
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-I" + CSRC,
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "--offload-device-only", "-S", "-o", str(asm),
           os.path.join(CSRC, "match.hip")]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    shapes = loop_shapes(asm.read_text().splitlines())
    # Modes 0..3 (count, trial, log, budgeted count), binary and merge probes.
    assert len(shapes) >= 6, shapes
    bad = {k: v for k, v in shapes.items() if v != (1, 1)}
    assert not bad, f"task fetch and broadcast not in one outermost loop: {bad}"


def test_loop_shapes_flags_a_nested_broadcast():
    """The checker itself: a broadcast inside a nested loop is reported."""
    good = ["_ZN12_GLOBAL__N_112match_kernelILi0EEEvv:",
            ".LBB0_1:  ; =>This Loop Header: Depth=1",
            "  global_atomic_add v1, v2, v3, s[4:5] sc0",
            "  ds_bpermute_b32 v0, v1, v2",
            ".Lfunc_end0:"]
    bad = ["_ZN12_GLOBAL__N_112match_kernelILi2EEEvv:",
           ".LBB1_1:  ; =>This Loop Header: Depth=1",
           "  global_atomic_add v1, v2, v3, s[4:5] sc0",
           ".LBB1_4:  ;   =>  This Loop Header: Depth=2",
           "  ds_bpermute_b32 v0, v1, v2",
           ".Lfunc_end1:"]
    shapes = loop_shapes(good + bad)
    assert shapes["_ZN12_GLOBAL__N_112match_kernelILi0EEEvv"] == (1, 1)
    assert shapes["_ZN12_GLOBAL__N_112match_kernelILi2EEEvv"] == (1, 2)
