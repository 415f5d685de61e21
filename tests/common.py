"""Shared helpers of the tests: the reference's sample reads, its golden
consensus strings and invariant checks, and the product driven through the
same checker harness as the oracle."""
import gzip
import json
import os

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)
PRODUCT_SO = os.path.join(ROOT, "spoa_amd", "libspoa_amd.so")

# The 9 scoring modes of the reference's golden tests
# (test/spoa_test.cpp:154-440): {SW, NW, OV} x {linear, affine, convex}.
MODES = []
for _t in (ol.SW, ol.NW, ol.OV):
    for _p in ((5, -4, -8, -8, -8, -8), (5, -4, -8, -6, -8, -6), (5, -4, -8, -6, -10, -2)):
        MODES.append((_t,) + _p)


def sample_reads():
    """(sequence, quality) of the 55 reads of the reference's test data."""
    lines = gzip.open(os.path.join(GOLDEN, "sample.fastq.gz"), "rt").read().split("\n")
    out = []
    for i in range(0, len(lines) - 3, 4):
        if lines[i].startswith("@"):
            out.append((lines[i + 1], lines[i + 3]))
    return out


def golden_consensus():
    with open(os.path.join(GOLDEN, "reference_consensus.json")) as f:
        return json.load(f)


def check_msa(msa, reads):
    """The reference's Check() invariants (test/spoa_test.cpp:56-78)."""
    assert len(msa) == len(reads)
    rs = len(msa[0])
    gaps = [0] * rs
    for row in msa:
        assert len(row) == rs
        for i, ch in enumerate(row):
            gaps[i] += ch == "-"
    assert all(g < len(msa) for g in gaps)
    for row, r in zip(msa, reads):
        assert row.replace("-", "") == r


def product():
    """The product's C entry points driven by the checker harness."""
    return ol.Checker(PRODUCT_SO, "spoa_")
