"""The device-resident window graph's algorithm (spoa_amd/csrc/window_graph.hpp:
AddAlignment, TopologicalSort, the alignment kernels' graph block, consensus
and MSA) compiled for the CPU with a one-lane team and checked, after every
read of 540 seeded windows in all nine scoring modes (with and without
Phred weights, with lower-case and N bases, with unaligned prefixes), against
the product's host graph (PoaGraph, itself pinned to the oracle) and the
oracle: ranks, node ids, predecessor lists, sink flags, first columns by the
reference's recurrence, consensus and MSA.  The 16-bit sort space is run
with a three-entry stack so that its overflow path (a retry in 32-bit
device scratch) is exercised too, and every other read the sort runs again
on a 4-lane team of threads in lockstep (its lists then span several lane
chunks), which must give the same ranks and graph block."""
import os
import subprocess

import common


def test_window_graph_against_host_graph_and_oracle(tmp_path):
    exe = tmp_path / "window_graph_test"
    root = common.ROOT
    subprocess.check_call([
        "g++", "-std=c++17", "-O2", "-pthread", "-I", os.path.join(root, "include"), "-I", os.path.join(root, "spoa_amd", "csrc"),
        os.path.join(root, "tests", "cpp", "window_graph_test.cpp"),
        os.path.join(root, "spoa_amd", "csrc", "poa_graph.cpp"),
        os.path.join(root, "oracle", "spoa_oracle.cpp"), "-o", str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.startswith("ok 540 windows")
