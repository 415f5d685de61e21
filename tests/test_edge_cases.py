"""Edge cases of the reference's own tests and error paths, on both sides.

* The reference's API edge tests (test/spoa_test.cpp:96-129): EmptyInput,
  LargeInput (Prealloc), Clear.
* The conditions Align and Prealloc report
  (src/simd_alignment_engine_implementation.hpp:381-422, :703-780):
  "too large sequence!", "possible overflow!" on either side of its bound,
  "insufficient memory!" under a device memory budget.
* Local alignments the reference scores at int32 whose scores fit int16 run
  on int16 rows while m * (n + 8) <= 30000: a near-perfect read on either
  side of that edge.
* The strand-ambiguous caller's reverse complement (src/main.cpp:286, 296):
  one rule in the product, the oracle, the reference shim and the harness,
  checked on every byte value and on windows of lower-case and IUPAC reads.

CPU tests drive the oracle and the compiled reference; GPU tests drive the
product through its C ABIs (spoa_capi.h and spoa_hip.h)."""
import ctypes
import os
import random
import subprocess

import pytest

import common
import oracle_lib as ol

# the reference's Align bound for Create(kSW, 1, -1, -128) against a
# one-node graph: WorstCaseAlignmentScore(len + 8, 1) = -128 * (len + 9),
# below INT32_MIN + 1024 = -128 * 16777208 from this length on
OVERFLOW_LEN = 16777200
OVERFLOW_PARAMS = (ol.SW, 1, -1, -128)

PRE_TOO_LARGE = "[spoa::SimdAlignmentEngine::Prealloc] error: too large sequence!"
PRE_NO_MEMORY = "[spoa::SimdAlignmentEngine::Prealloc] error: insufficient memory!"
ALIGN_OVERFLOW = "[spoa::SimdAlignmentEngine::Align] error: possible overflow!"
ALIGN_NO_MEMORY = "[spoa::SimdAlignmentEngine::Align] error: insufficient memory!"


def _rand_seq(rng, n, alpha="ACGT"):
    return "".join(rng.choice(alpha) for _ in range(n))


def _one_node_graph(chk, base="A"):
    """A graph holding one read of one base (the first read of a window is
    added with an empty alignment)."""
    gr = chk.graph()
    chk.add_alignment(gr, [], base)
    return gr


def _strand_windows(seed):
    """Windows whose reads mix lower-case bases, IUPAC codes and N, half of
    them reverse-complemented (by the shared rule)."""
    rng = random.Random(seed)
    windows = []
    for _ in range(3):
        templ = _rand_seq(rng, rng.randint(120, 220))
        reads = []
        for k in range(rng.randint(5, 8)):
            out = []
            for b in templ:
                x = rng.random()
                if x < 0.04:
                    out.append(rng.choice("ACGT"))
                elif x < 0.07:
                    continue
                elif x < 0.10:
                    out.append(b + rng.choice("ACGT"))
                elif x < 0.13:
                    out.append(rng.choice("RYKMSWBDHVN"))
                else:
                    out.append(b)
            r = "".join(out)
            if k % 3 == 1:
                r = r.lower()
            elif k % 3 == 2:
                r = "".join(c.lower() if rng.random() < 0.3 else c for c in r)
            if k % 2 == 1:
                r = ol.revcomp(r)
            reads.append(r)
        windows.append(reads)
    return windows


# ----------------------------------------------------------------- CPU --

def _lib_revcomp(lib, name, data):
    fn = getattr(lib, name)
    fn.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p]
    out = ctypes.create_string_buffer(len(data))
    assert fn(data, len(data), out) == len(data)
    return out.raw[:len(data)]


def test_reverse_complement_rule_is_one_rule():
    """Product, oracle, reference shim and harness agree on every byte."""
    data = bytes(range(1, 256))
    want = ol.revcomp(data)
    assert _lib_revcomp(ol.oracle().lib, "oracle_reverse_complement", data) == want
    ref = ol.reference()
    if ref is not None:
        assert _lib_revcomp(ref.lib, "ref_reverse_complement", data) == want
    if os.path.exists(common.PRODUCT_SO):
        assert _lib_revcomp(ctypes.CDLL(common.PRODUCT_SO), "spoa_reverse_complement", data) == want


def test_reverse_complement_closed_form():
    """Watson-Crick pairs, reversal, involution on upper-case codes, case
    folded, other bytes kept."""
    assert ol.revcomp("AACGTT") == "AACGTT"
    assert ol.revcomp("AAACG") == "CGTTT"
    assert ol.revcomp("acgu") == "ACGT"
    assert ol.revcomp("RYKMBVDHSW") == "WSDHBVKMRY"
    codes = "ACGTRYKMBVDHSWN-"
    assert ol.revcomp(ol.revcomp(codes)) == codes
    assert ol.revcomp("n-*x") == "x*-n"
    assert ol.revcomp("") == ""


def test_strand_windows_of_mixed_case_and_iupac_reads():
    """Strand mode on lower-case / IUPAC reads: the oracle equals the
    compiled reference call by call (both with the shared complement), and
    the oracle's own C window runner (its CPU-baseline path) gives the same
    consensus as the harness."""
    orc = ol.oracle()
    ref = ol.reference()
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    for reads in _strand_windows(7):
        co, mo, ro = ol.run_window(orc, params, reads, strand=True)
        if ref is not None:
            cr, mr, rr = ol.run_window(ref, params, reads, strand=True)
            assert co == cr and mo == mr
            for a, b in zip(ro, rr):
                assert a["pairs"] == b["pairs"] and a["pairs_rc"] == b["pairs_rc"]
                assert a["kept_forward"] == b["kept_forward"]
        seqs = "".join(reads).encode()
        offs = [0]
        for r in reads:
            offs.append(offs[-1] + len(r))
        fn = orc.lib.oracle_run_window
        fn.argtypes = [ctypes.c_int] * 7 + [ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p,
                                            ctypes.c_int, ctypes.c_char_p, ctypes.c_uint32]
        out = ctypes.create_string_buffer(1 << 16)
        n = fn(*params, len(reads), (ctypes.c_uint64 * len(offs))(*offs), seqs, 1, out, 1 << 16)
        assert n >= 0 and out.raw[:n].decode() == co


def test_checkers_empty_input():
    """EmptyInput (test/spoa_test.cpp:96-101) on the oracle and reference."""
    for chk in filter(None, [ol.oracle(), ol.reference()]):
        eng = chk.engine(ol.SW, 1, -1, -1)
        gr = chk.graph()
        assert chk.align(eng, gr, "") == ([], None, 0)
        chk.add_alignment(gr, [], "ACGT")
        assert chk.align(eng, gr, "") == ([], None, 0)
        chk.free_graph(gr)
        chk.free_engine(eng)


def test_checkers_prealloc_and_overflow():
    """LargeInput (test/spoa_test.cpp:103-119) and the overflow bound, on
    the oracle and the compiled reference."""
    orc = ol.oracle()
    eng = orc.engine(ol.SW, 1, -1, -1)
    f = orc.lib.oracle_engine_prealloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    assert f(eng, 0xFFFFFFFF, 1) == -1 and orc.error() == PRE_TOO_LARGE
    assert f(eng, (1 << 31) - 1, 255) == 0  # bound past int32: nothing sized
    assert f(eng, 1000, 4) == 16
    orc.free_engine(eng)
    ref = ol.reference()
    if ref is not None:
        e = ref.engine(ol.SW, 1, -1, -1)
        g = ref.lib.ref_engine_prealloc
        g.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
        assert g(e, 0xFFFFFFFF, 1) == -1 and ref.error() == PRE_TOO_LARGE
        assert g(e, (1 << 31) - 1, 255) == 0
        ref.free_engine(e)
    rng = random.Random(3)
    long_read = _rand_seq(rng, OVERFLOW_LEN)
    for chk in filter(None, [orc, ref]):
        eng = chk.engine(*OVERFLOW_PARAMS)
        gr = _one_node_graph(chk)
        with pytest.raises(ValueError) as ex:
            chk.align(eng, gr, long_read)
        assert str(ex.value) == ALIGN_OVERFLOW
        chk.free_graph(gr)
        chk.free_engine(eng)


def test_checkers_at_the_overflow_bound():
    """One base short of the bound (the bound itself) the reference still
    aligns, at int32, and the oracle agrees with it (the reference's width
    is not observable through its API)."""
    rng = random.Random(4)
    read = _rand_seq(rng, OVERFLOW_LEN - 1)
    out = []
    for chk in filter(None, [ol.oracle(), ol.reference()]):
        eng = chk.engine(*OVERFLOW_PARAMS)
        gr = _one_node_graph(chk, read[5])
        out.append(chk.align(eng, gr, read))
        chk.free_graph(gr)
        chk.free_engine(eng)
    assert out[0][2] == 32 and out[0][1] == 1
    assert all(o[:2] == out[0][:2] for o in out)


def test_pass_splitter(tmp_path):
    """The batched form's pass splitter (spoa_amd/csrc/pass_split.hpp): the
    memory budget and the 32-bit offsets of a pass."""
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <cassert>
#include <cstdio>
#include "pass_split.hpp"
using spoa_amd::PassItem;
using spoa_amd::PassEnd;
using spoa_amd::PassItemFits;
int main() {
  const std::uint64_t M = 0xFFFFFFFFull;
  PassItem a[6] = {{40, 10, 10, 10}, {40, 10, 10, 10}, {30, 10, 10, 10},
                   {90, 10, 10, 10}, {5, 10, 10, 10}, {5, 10, 10, 10}};
  assert(PassEnd(a, 0, 6, 100) == 2);   // 40 + 40 + 30 > 100
  assert(PassEnd(a, 2, 6, 100) == 3);   // 30 + 90 > 100
  assert(PassEnd(a, 3, 6, 100) == 6);   // 90 + 5 + 5
  assert(PassEnd(a, 6, 6, 100) == 6);
  assert(PassEnd(a, 0, 6, 1000) == 6);
  // 32-bit totals: graph words, read bytes, pair slots
  PassItem w[3] = {{1, M / 2, 1, 1}, {1, M / 2, 1, 1}, {1, 2, 1, 1}};
  assert(PassEnd(w, 0, 3, 1000) == 2 && PassEnd(w, 2, 3, 1000) == 3);
  PassItem s[2] = {{1, 1, M, 1}, {1, 1, 1, 1}};
  assert(PassEnd(s, 0, 2, 1000) == 1);
  PassItem p[3] = {{1, 1, 1, M - 5}, {1, 1, 1, 5}, {1, 1, 1, 1}};
  assert(PassEnd(p, 0, 3, 1000) == 2 && PassEnd(p, 2, 3, 1000) == 3);
  // one job alone
  assert(PassItemFits({100, M, M, M}, 100));
  assert(!PassItemFits({101, 1, 1, 1}, 100));
  assert(!PassItemFits({1, M + 1, 1, 1}, 100));
  assert(!PassItemFits({1, 1, M + 1, 1}, 100));
  assert(!PassItemFits({1, 1, 1, M + 1}, 100));
  // a pass always takes its first job
  PassItem big[2] = {{100, 1, 1, 1}, {1, 1, 1, 1}};
  assert(PassEnd(big, 0, 2, 100) == 1);
  std::puts("ok");
  return 0;
}
''')
    exe = tmp_path / "t"
    inc = os.path.join(common.ROOT, "spoa_amd", "csrc")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", inc, str(src), "-o", str(exe)])
    assert subprocess.check_output([str(exe)]).strip() == b"ok"


# ----------------------------------------------------------------- GPU --

class HipGraph(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint32), ("row_char", ctypes.c_void_p), ("node_id", ctypes.c_void_p),
                ("pred_off", ctypes.c_void_p), ("pred_rank", ctypes.c_void_p), ("is_sink", ctypes.c_void_p)]


class HipParams(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("subtype", ctypes.c_int32), ("m", ctypes.c_int8),
                ("n", ctypes.c_int8), ("g", ctypes.c_int8), ("e", ctypes.c_int8), ("q", ctypes.c_int8),
                ("c", ctypes.c_int8)]


def _hip_lib():
    import spoa_amd
    L = spoa_amd.lib()
    L.spoa_hip_engine_create.argtypes = [ctypes.c_int, ctypes.POINTER(HipParams), ctypes.POINTER(ctypes.c_void_p)]
    L.spoa_hip_engine_destroy.argtypes = [ctypes.c_void_p]
    L.spoa_hip_align.argtypes = [ctypes.c_void_p, ctypes.POINTER(HipGraph), ctypes.c_char_p, ctypes.c_uint32,
                                 ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.spoa_hip_prealloc.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    return L


@pytest.mark.gpu
def test_empty_input_through_both_c_abis():
    """EmptyInput: Align("") returns an empty alignment and leaves the score
    unwritten, on an empty and on a non-empty graph; through spoa_capi and
    through spoa_hip_align with a one-node rank view."""
    prod = common.product()
    eng = prod.engine(ol.SW, 1, -1, -1)
    gr = prod.graph()
    assert prod.align(eng, gr, "") == ([], None, 0)
    prod.add_alignment(gr, [], "ACGT")
    assert prod.align(eng, gr, "") == ([], None, 0)
    assert prod.align(eng, gr, "ACGT")[0] == [(0, 0), (1, 1), (2, 2), (3, 3)]
    prod.free_graph(gr)
    prod.free_engine(eng)

    L = _hip_lib()
    h = ctypes.c_void_p()
    prm = HipParams(0, 0, 1, -1, -1, -1, -1, -1)
    assert L.spoa_hip_engine_create(0, ctypes.byref(prm), ctypes.byref(h)) == 0
    try:
        row_char = (ctypes.c_uint8 * 1)(ord("A"))
        node_id = (ctypes.c_int32 * 1)(0)
        pred_off = (ctypes.c_uint32 * 2)(0, 0)
        pred_rank = (ctypes.c_uint32 * 1)(0)
        sink = (ctypes.c_uint8 * 1)(1)
        pairs = (ctypes.c_int32 * 16)()
        for v in (0, 1):
            g = HipGraph(v, ctypes.cast(row_char, ctypes.c_void_p), ctypes.cast(node_id, ctypes.c_void_p),
                         ctypes.cast(pred_off, ctypes.c_void_p), ctypes.cast(pred_rank, ctypes.c_void_p),
                         ctypes.cast(sink, ctypes.c_void_p))
            for seq in (b"", b"A"):
                npairs, score, has = ctypes.c_uint32(99), ctypes.c_int32(-7), ctypes.c_int32(5)
                rc = L.spoa_hip_align(h, ctypes.byref(g), seq, len(seq), pairs, 8, ctypes.byref(npairs),
                                      ctypes.byref(score), ctypes.byref(has))
                assert rc == 0
                if v == 0 or not seq:
                    assert (npairs.value, has.value, score.value) == (0, 0, -7)
                else:
                    assert (npairs.value, has.value, score.value) == (1, 1, 1)
                    assert list(pairs[:2]) == [0, 0]
    finally:
        L.spoa_hip_engine_destroy(h)


@pytest.mark.gpu
def test_large_input_prealloc():
    """LargeInput (test/spoa_test.cpp:103-119): the reference's message from
    character 11 on; a bound past int32 sizes nothing and is no error."""
    import spoa_amd
    eng = spoa_amd.AlignmentEngine.create(spoa_amd.SW, 1, -1, -1)
    with pytest.raises(ValueError) as ex:
        eng.prealloc(0xFFFFFFFF, 1)
    assert str(ex.value) == PRE_TOO_LARGE
    assert str(ex.value)[11:] == "AlignmentEngine::Prealloc] error: too large sequence!"
    eng.prealloc((1 << 31) - 1, 255)
    eng.prealloc(1000, 4)
    L = _hip_lib()
    h = ctypes.c_void_p()
    prm = HipParams(0, 0, 1, -1, -1, -1, -1, -1)
    assert L.spoa_hip_engine_create(0, ctypes.byref(prm), ctypes.byref(h)) == 0
    try:
        assert L.spoa_hip_prealloc(h, 0xFFFFFFFF, 1) == 2  # SPOA_HIP_TOO_LARGE
        assert L.spoa_hip_last_error().decode() == PRE_TOO_LARGE
        assert L.spoa_hip_prealloc(h, (1 << 31) - 1, 255) == 0
    finally:
        L.spoa_hip_engine_destroy(h)


@pytest.mark.gpu
def test_clear_then_realign_reproduces_consensus():
    """Clear (test/spoa_test.cpp:121-129): SW linear on the reference's
    sample reads, Graph::Clear, the same reads again."""
    import spoa_amd
    params = common.MODES[0]
    reads = [s for s, _ in common.sample_reads()]
    eng = spoa_amd.AlignmentEngine.create(*params)
    eng.prealloc(max(len(r) for r in reads), 4)
    gr = spoa_amd.Graph()
    for r in reads:
        a, _ = eng.align(r, gr)
        gr.add_alignment(a, r)
    c = gr.generate_consensus()
    gr.clear()
    assert gr.num_nodes == 0
    for r in reads:
        a, _ = eng.align(r, gr)
        gr.add_alignment(a, r)
    assert gr.generate_consensus() == c
    golden = [g for g in common.golden_consensus().values() if g["params"] == list(params) and not g["quality"]]
    assert c == golden[0]["consensus"]


@pytest.mark.gpu
def test_possible_overflow_and_its_bound():
    """"possible overflow!" for a read one base past the bound, before any
    device work; one base short of it the product aligns as the oracle
    does (int32, a 16.7M-column row in 4,096 tiles)."""
    prod = common.product()
    orc = ol.oracle()
    rng = random.Random(4)
    read = _rand_seq(rng, OVERFLOW_LEN)
    eng = prod.engine(*OVERFLOW_PARAMS)
    gr = _one_node_graph(prod)
    with pytest.raises(ValueError) as ex:
        prod.align(eng, gr, read)
    assert str(ex.value) == ALIGN_OVERFLOW
    prod.free_graph(gr)
    short = read[:-1]
    gr = _one_node_graph(prod, short[5])
    got = prod.align(eng, gr, short)
    prod.free_graph(gr)
    prod.free_engine(eng)
    oe = orc.engine(*OVERFLOW_PARAMS)
    og = _one_node_graph(orc, short[5])
    want = orc.align(oe, og, short)
    orc.free_graph(og)
    orc.free_engine(oe)
    assert got == want and got[2] == 32


@pytest.mark.gpu
def test_insufficient_memory_under_a_budget():
    """A device memory budget below one job's DP matrices: Align and
    Prealloc report "insufficient memory!"; with the budget raised the same
    engine aligns as the oracle.  Prealloc keeps the engine's matrices
    within the budget."""
    import spoa_amd
    L = spoa_amd.lib()
    rng = random.Random(11)
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    eng = spoa_amd.AlignmentEngine.create(*params)
    dev = eng.device
    gr = spoa_amd.Graph()
    first = _rand_seq(rng, 900)
    gr.add_alignment([], first)
    read = first[:300] + _rand_seq(rng, 20) + first[320:]
    budget0 = L.spoa_hip_memory_budget(dev)
    need = L.spoa_hip_job_footprint(dev, gr.num_nodes, len(read), 0)
    assert need > 0
    assert L.spoa_hip_set_memory_budget(dev, need - 1) == 0
    with pytest.raises(ValueError) as ex:
        eng.align(read, gr)
    assert str(ex.value) == ALIGN_NO_MEMORY
    with pytest.raises(ValueError) as ex:
        eng.prealloc(5000, 4)
    assert str(ex.value) == PRE_NO_MEMORY
    assert L.spoa_hip_set_memory_budget(dev, need) == 0
    got = eng.align(read, gr)
    orc = ol.oracle()
    oe = orc.engine(*params)
    og = orc.graph()
    orc.add_alignment(og, [], first)
    want = orc.align(oe, og, read)
    orc.free_graph(og)
    orc.free_engine(oe)
    assert list(got[0]) == want[0] and got[1] == want[1]
    # growing Prealloc calls stay within a budget that holds each alone
    eng2 = spoa_amd.AlignmentEngine.create(*params)
    dev2 = eng2.device
    sizes = [L.spoa_hip_job_footprint(dev2, 4 * n, n, 0) for n in (600, 900, 1200)]
    budget = sizes[-1] + sizes[-1] // 4
    assert sizes[0] + sizes[1] + sizes[2] > budget
    assert L.spoa_hip_set_memory_budget(dev2, budget) == 0
    for n in (600, 900, 1200):
        eng2.prealloc(n, 4)
        assert L.spoa_hip_memory_in_use(dev2) <= budget
    assert L.spoa_hip_memory_in_use(dev2) >= sizes[-1]
    assert budget0 > 0


@pytest.mark.gpu
@pytest.mark.parametrize("n", [5992, 5993])
def test_local_int32_width_int16_rows_edge(n):
    """Local convex alignments the reference scores at int32 (width 32 from
    its bound) keep int16 rows while m * (n + 8) <= 30000: 5 * 6000 is the
    last length that does.  A near-perfect read scores close to m * n on
    either side of the edge; pairs and score equal the oracle's, and the
    DP footprint halves on the int16 side."""
    import spoa_amd
    params = (ol.SW, 5, -4, -8, -6, -10, -4)
    rng = random.Random(n)
    first = _rand_seq(rng, n)
    read = list(first)
    for k in (n // 3, 2 * n // 3):
        read[k] = "A" if read[k] != "A" else "C"
    read = "".join(read)
    prod = common.product()
    orc = ol.oracle()
    got = []
    for chk in (prod, orc):
        eng = chk.engine(*params)
        gr = chk.graph()
        chk.add_alignment(gr, [], first)
        got.append(chk.align(eng, gr, read))
        chk.free_graph(gr)
        chk.free_engine(eng)
    assert got[0] == got[1]
    assert got[0][2] == 32
    assert got[0][1] >= 5 * n - 40
    L = spoa_amd.lib()
    eng = spoa_amd.AlignmentEngine.create(*params)
    here = L.spoa_hip_job_footprint(eng.device, n, n, 0)
    before = L.spoa_hip_job_footprint(eng.device, n, n - 1, 0)  # int16 rows on both sides
    if n == 5992:
        assert here < 1.01 * before
    else:
        assert here > 1.9 * before
    # the batched form on the same window
    ws = spoa_amd.Windows.from_lists([[first, read]])
    res = spoa_amd.run_windows(ws, *params, mode="batched", msa=True)
    cons, msa, _ = ol.run_window(orc, params, [first, read])
    assert res.consensus[0] == cons and res.msa[0] == msa


@pytest.mark.gpu
def test_strand_mode_on_mixed_case_and_iupac_reads():
    """The batched and per-call forms in strand mode, on windows of
    lower-case and IUPAC reads, against the oracle (same complement)."""
    import spoa_amd
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    windows = _strand_windows(7)
    ws = spoa_amd.Windows.from_lists(windows)
    for mode in ("batched", "per_call"):
        res = spoa_amd.run_windows(ws, *params, mode=mode, strand=True, msa=True)
        for w, reads in enumerate(windows):
            cons, msa, recs = ol.run_window(ol.oracle(), params, reads, strand=True)
            assert res.consensus[w] == cons, (mode, w)
            assert res.msa[w] == msa, (mode, w)
            assert res.reversed[w] == [0 if r["kept_forward"] else 1 for r in recs], (mode, w)
