"""Device-resident window graphs (spoa_hip_windows_*): AddAlignment,
TopologicalSort, the alignment kernels' graph block, consensus and MSA on
the device.

* Ranks after every read: every window's rank_to_node, read back after each
  round, equals the oracle's after the same reads (all three alignment
  types, all three gap models, strand mode, Phred weights).
* Whole runs equal the compiled reference's vectors and the oracle
  (consensus, MSA, kept strands), and the host-merged batched form.
* A window too large for the LDS sort space, or whose walk outgrows the
  LDS stack, sorts in device memory."""
import ctypes
import gzip
import hashlib
import json
import os
import random

import pytest

import common
import oracle_lib as ol
import spoa_amd

pytestmark = pytest.mark.gpu


def _lib():
    L = spoa_amd.lib()
    vp = ctypes.c_void_p
    L.spoa_hip_windows_create.argtypes = [vp, ctypes.c_uint32, vp, vp, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.c_int32, ctypes.POINTER(vp)]
    L.spoa_hip_windows_destroy.argtypes = [vp]
    L.spoa_hip_windows_rounds.restype = ctypes.c_uint32
    L.spoa_hip_windows_rounds.argtypes = [vp]
    L.spoa_hip_windows_submit.argtypes = [vp, ctypes.c_uint32]
    L.spoa_hip_windows_wait.argtypes = [vp]
    L.spoa_hip_windows_ranks.argtypes = [vp, ctypes.c_uint32, vp, ctypes.c_uint32]
    L.spoa_hip_windows_last_error.restype = ctypes.c_char_p
    L.spoa_hip_windows_last_error.argtypes = [vp]
    return L


def _mutate(rng, templ, p, alpha="ACGT"):
    out = []
    for b in templ:
        x = rng.random()
        if x < p / 3:
            out.append(rng.choice(alpha))
        elif x < 2 * p / 3:
            out.append(b + rng.choice(alpha))
        elif x < p:
            continue
        else:
            out.append(b)
    return "".join(out) or templ


def _windows(seed, n, tlen=(80, 260), nreads=(3, 12), p=0.12, alpha="ACGT", revcomp=False):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        templ = "".join(rng.choice("ACGT") for _ in range(rng.randint(*tlen)))
        reads = []
        for k in range(rng.randint(*nreads)):
            r = _mutate(rng, templ, p, alpha)
            if revcomp and k % 2 == 1:
                r = ol.revcomp(r)
            reads.append(r)
        out.append(reads)
    return out


MODES = [(ol.NW, 3, -5, -3), (ol.OV, 5, -4, -8, -6), (ol.SW, 5, -4, -8, -6, -10, -4), (ol.NW, 5, -4, -8, -6, -10, -4),
         (ol.SW, 5, -4, -8), (ol.OV, 5, -4, -8, -6, -10, -2)]


@pytest.mark.parametrize("params", MODES)
def test_ranks_after_every_read(params):
    """rank_to_node of every window after every round equals the oracle's
    after the same reads (windows of uneven read counts and lengths)."""
    L = _lib()
    strand = params == MODES[3]
    windows = _windows(hash(params) % 1000, 6, revcomp=strand)
    ws = spoa_amd.Windows.from_lists(windows)
    eng = spoa_amd.AlignmentEngine.create(*params)
    h = ctypes.c_void_p()
    assert L.spoa_hip_windows_create(eng.device, ws.n_windows, ws.win_read_off.ctypes.data, ws.read_off.ctypes.data,
                                     ws.seqs, None, int(strand), ctypes.byref(h)) == 0, L.spoa_hip_last_error()
    orc = ol.oracle()
    oe = orc.engine(*params)
    ogs = [orc.graph() for _ in windows]
    try:
        rounds = L.spoa_hip_windows_rounds(h)
        assert rounds == max(len(w) for w in windows)
        for r in range(rounds):
            assert L.spoa_hip_windows_submit(h, r) == 0, L.spoa_hip_windows_last_error(h)
            assert L.spoa_hip_windows_wait(h) == 0, L.spoa_hip_windows_last_error(h)
            for w, reads in enumerate(windows):
                if r >= len(reads):
                    continue
                read = reads[r]
                pairs, score, _ = orc.align(oe, ogs[w], read)
                seq = read
                if strand:
                    rc = ol.revcomp(read)
                    pr, sr, _ = orc.align(oe, ogs[w], rc)
                    if not ((score or 0) >= (sr or 0)):
                        pairs, seq = pr, rc
                orc.add_alignment(ogs[w], pairs, seq)
                want = orc.ranks(ogs[w])
                buf = (ctypes.c_uint32 * max(1, len(want) + 8))()
                n = L.spoa_hip_windows_ranks(h, w, buf, len(want) + 8)
                assert n == len(want), (r, w, n)
                assert list(buf[:n]) == want, (r, w)
    finally:
        L.spoa_hip_windows_destroy(h)
        for g in ogs:
            orc.free_graph(g)
        orc.free_engine(oe)


@pytest.fixture(scope="module")
def vectors():
    with gzip.open(os.path.join(common.GOLDEN, "ref_vectors.json.gz"), "rt") as f:
        return json.load(f)


def _sha(rows):
    return hashlib.sha1("\n".join(rows).encode()).hexdigest()


def test_resident_matches_reference_vectors(vectors):
    """Config-shaped windows recorded from the compiled reference, the
    strand-ambiguous long window included."""
    for cfg in (1, 2, 3, 4):
        recs = [r for r in vectors["configs"] if r["config"] == cfg]
        ws = spoa_amd.Windows.from_lists([r["reads"] for r in recs])
        strand = cfg == 4
        res = spoa_amd.run_windows(ws, *recs[0]["params"], mode="resident", strand=strand, msa=True)
        for w, r in enumerate(recs):
            assert res.consensus[w] == r["consensus"], (cfg, w)
            assert _sha(res.msa[w]) == r["msa_sha1"], (cfg, w)
            if strand:
                assert res.reversed[w] == [0 if c["kept_forward"] else 1 for c in r["calls"]]


@pytest.mark.parametrize("mode", range(9))
def test_resident_reference_sample_with_weights(mode):
    """The reference's sample reads with Phred weights in all nine modes
    (test/spoa_test.cpp:154-440): golden consensus, oracle MSA."""
    params = common.MODES[mode]
    data = common.sample_reads()
    reads = [s for s, _ in data]
    quals = [q for _, q in data]
    ws = spoa_amd.Windows.from_lists([reads], [quals])
    res = spoa_amd.run_windows(ws, *params, mode="resident", msa=True)
    golden = [g for g in common.golden_consensus().values() if g["params"] == list(params) and g["quality"]]
    assert res.consensus[0] == golden[0]["consensus"]
    weights = [[ord(c) - 33 for c in q] for q in quals]
    _, msa, _ = ol.run_window(ol.oracle(), params, reads, weights)
    assert res.msa[0] == msa


@pytest.mark.parametrize("cfg", [1, 3])
def test_resident_equals_oracle_and_host_merged_form(cfg):
    c = spoa_amd.CONFIGS[cfg]
    ws = spoa_amd.config_windows(cfg, seed=777, n_windows=24)
    res = spoa_amd.run_windows(ws, *c["params"], mode="resident", msa=True)
    host = spoa_amd.run_windows(ws, *c["params"], mode="batched", msa=True)
    assert res.consensus == host.consensus
    assert res.msa == host.msa
    for w in range(0, ws.n_windows, 3):
        cons, msa, _ = ol.run_window(ol.oracle(), c["params"], ws.reads(w))
        assert res.consensus[w] == cons and res.msa[w] == msa, w


def test_resident_mixed_alphabet_and_lowercase():
    """Windows with N, IUPAC and lower-case bases (groups of up to eight
    codes, codes in order of first appearance)."""
    params = (ol.OV, 5, -4, -8, -6, -10, -4)
    windows = _windows(5, 5, p=0.2, alpha="ACGTNRYacgt")
    res = spoa_amd.run_windows(windows, *params, mode="resident", msa=True)
    for w, reads in enumerate(windows):
        cons, msa, _ = ol.run_window(ol.oracle(), params, reads)
        assert res.consensus[w] == cons and res.msa[w] == msa, w


def test_resident_large_window_sorts_in_device_memory():
    """A 10 kb, 24-read window grows past the LDS sort space (tens of
    thousands of nodes): the sort runs in the window's device scratch, and
    the results equal the host-merged batched form."""
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    windows = _windows(9, 1, tlen=(10000, 10000), nreads=(24, 24), p=0.10)
    host = spoa_amd.run_windows(windows, *params, mode="batched", msa=True)
    res = spoa_amd.run_windows(windows, *params, mode="resident", msa=True)
    assert res.consensus == host.consensus
    assert res.msa == host.msa


def test_resident_deep_stack_sorts_in_device_memory():
    """An overlap read with a 1,500-base overhang becomes a chain of new
    nodes that the depth-first walk descends in one go; the later reads are
    50 bases, so the LDS room of their rounds (sized by the read) is smaller
    than that stack: the 16-bit sort reports the overflow and the window
    sorts again in its device scratch.  Consensus and MSA equal the
    oracle's; a neighbouring window without the chain too."""
    rng = random.Random(5)
    templ = "".join(rng.choice("ACGT") for _ in range(300))
    over = "".join(rng.choice("ACGT") for _ in range(1500))
    deep = [templ, over + templ] + [templ[i:i + 50] for i in (10, 100, 170, 240)]
    plain = [_mutate(rng, templ, 0.1) for _ in range(6)]
    params = (ol.OV, 5, -4, -8, -6, -10, -4)
    res = spoa_amd.run_windows([deep, plain], *params, mode="resident", msa=True)
    for w, reads in enumerate((deep, plain)):
        cons, msa, _ = ol.run_window(ol.oracle(), params, reads)
        assert res.consensus[w] == cons and res.msa[w] == msa, w


def test_context_reuses_buffers_across_window_sets():
    """One resident context runs a larger set, a smaller and different one
    (strand mode, other read lengths), then the larger one again: each run
    takes over the buffers the last one left and equals the oracle run
    alone on every window (and the first run, for the repeat)."""
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    big = _windows(21, 12, tlen=(300, 700), nreads=(5, 14), p=0.12, revcomp=True)
    small = _windows(22, 5, tlen=(60, 200), nreads=(2, 6), p=0.15, revcomp=True)
    ctx = spoa_amd.BatchContext(*params, strand=True, msa=True, resident=True)
    first = ctx.run(big)
    second = ctx.run(small)
    third = ctx.run(big)
    for windows, res in ((big, first), (small, second)):
        for w, reads in enumerate(windows):
            cons, msa, _ = ol.run_window(ol.oracle(), params, reads, strand=True)
            assert res.consensus[w] == cons and res.msa[w] == msa, w
    assert list(third.consensus) == list(first.consensus)
    assert [third.msa[w] for w in range(len(big))] == [first.msa[w] for w in range(len(big))]
    ctx.close()
