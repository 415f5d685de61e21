"""GPU parity of engines that stand in for the reference's scalar engine
(SisdAlignmentEngine, src/sisd_alignment_engine.cpp; spoa_sisd_engine_new,
spoa_hip_engine_create_sisd, SPOA_HIP_ENGINE=sisd): the HIP path against
the oracle's restatement of that engine (pinned to the compiled scalar
engine by tests/test_oracle_sisd.py) and against vectors recorded from the
compiled scalar engine -- alignment pairs, scores, reported width (32),
consensus and MSA, in the per-call, batched and device-resident forms."""
import gzip
import json
import os
import random

import pytest

import common
import oracle_lib as ol
import spoa_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    return common.product()


@pytest.fixture(scope="module")
def orc():
    return ol.oracle()


@pytest.fixture(scope="module")
def vectors():
    with gzip.open(os.path.join(common.GOLDEN, "ref_sisd_vectors.json.gz"), "rt") as f:
        return json.load(f)


def _compare(prod, orc, params, reads, weights=None, strand=False):
    cp, mp, rp = ol.run_window(prod, params, reads, weights, strand=strand, sisd=True)
    co, mo, ro = ol.run_window(orc, params, reads, weights, strand=strand, sisd=True)
    for k, (a, b) in enumerate(zip(rp, ro)):
        assert a["score"] == b["score"], ("score", params, k)
        assert a["pairs"] == b["pairs"], ("pairs", params, k)
        assert a["width"] == b["width"], ("width", params, k)
        if strand:
            assert a["pairs_rc"] == b["pairs_rc"], ("pairs_rc", params, k)
            assert a["score_rc"] == b["score_rc"], ("score_rc", params, k)
    assert cp == co
    assert mp == mo
    return cp, mp


def _mutate(rng, s, p, alpha):
    out = []
    for b in s:
        x = rng.random()
        if x < p / 3:
            out.append(rng.choice(alpha))
        elif x < 2 * p / 3:
            out.append(b + rng.choice(alpha))
        elif x >= p:
            out.append(b)
    return "".join(out) or rng.choice(alpha)


@pytest.mark.parametrize("quality", [False, True])
@pytest.mark.parametrize("mode", range(9))
def test_sisd_sample_modes(gpu, orc, mode, quality):
    """The reference's 18 end-to-end cases (test/spoa_test.cpp:154-440)
    with the scalar engine's alignments: golden consensus, oracle pairs."""
    params = common.MODES[mode]
    data = common.sample_reads()
    reads = [s for s, _ in data]
    weights = [[ord(c) - 33 for c in q] for _, q in data] if quality else None
    cons, msa = _compare(gpu, orc, params, reads, weights)
    golden = [g for g in common.golden_consensus().values()
              if g["params"] == list(params) and g["quality"] == quality]
    assert cons == golden[0]["consensus"]
    common.check_msa(msa, reads)


def _check_vectors(prod, rec, reads, weights=None, full=False):
    from test_oracle import msa_hash, pairs_hash
    cons, msa, recs = ol.run_window(prod, rec["params"], reads, weights, strand=rec["strand"], sisd=True)
    for k, (got, exp) in enumerate(zip(recs, rec["calls"])):
        assert got["score"] == exp["score"], k
        assert pairs_hash(got["pairs"]) == exp["sha1"], k
        if full:
            assert [list(p) for p in got["pairs"]] == exp["pairs"], k
        if rec["strand"]:
            assert got["score_rc"] == exp["score_rc"] and pairs_hash(got["pairs_rc"]) == exp["sha1_rc"], k
    assert cons == rec["consensus"]
    assert msa_hash(msa) == rec["msa_sha1"]


def test_sisd_reference_vectors(gpu, vectors):
    """Vectors recorded from the compiled scalar engine: its sample runs,
    benchmark-shaped windows (the strand-ambiguous long window included) and
    random windows over random scores and alphabets."""
    data = common.sample_reads()
    reads = [s for s, _ in data]
    quals = [[ord(c) - 33 for c in q] for _, q in data]
    for rec in vectors["sample"]:
        _check_vectors(gpu, rec, reads, quals if rec["quality"] else None)
    for rec in vectors["configs"]:
        print("config", rec["config"], rec["window"], flush=True)
        _check_vectors(gpu, rec, rec["reads"])
    for rec in vectors["random"]:
        if "error" in rec:
            with pytest.raises(ValueError) as ex:
                ol.run_window(gpu, rec["params"], rec["reads"], strand=rec["strand"], sisd=True)
            assert str(ex.value) == rec["error"]
        else:
            _check_vectors(gpu, rec, rec["reads"], full=True)


@pytest.mark.parametrize("seed", range(4))
def test_sisd_random_windows(gpu, orc, seed):
    """Seeded random windows over all 9 modes and random scores (some
    forcing int32 rows), short and ragged reads, odd alphabets, strands."""
    rng = random.Random(3000 + seed)
    for trial in range(12):
        if rng.random() < 0.3:
            params = (rng.randrange(3), rng.randint(0, 120), rng.randint(-120, 0),
                      rng.randint(-128, 0), rng.randint(-128, 0), rng.randint(-128, 0),
                      rng.randint(-128, 0))
        else:
            params = (rng.randrange(3), rng.randint(0, 10), rng.randint(-10, 0), rng.randint(-12, 0),
                      rng.randint(-12, 0), rng.randint(-14, 0), rng.randint(-12, 0))
        alpha = rng.choice(["ACGT", "AC", "ACGTN", "A"])
        tl = rng.choice([1, 5, 30, 120, 400, 1100])
        templ = "".join(rng.choice(alpha) for _ in range(tl))
        reads = []
        for _ in range(rng.randint(1, 6)):
            a = rng.randrange(tl)
            b = rng.randint(a + 1, tl)
            src = templ[a:b] if rng.random() < 0.5 else templ
            reads.append(_mutate(rng, src, rng.random() * 0.3, alpha))
        strand = rng.random() < 0.25
        try:
            ol.run_window(orc, params, reads, strand=strand, sisd=True)
        except ValueError as ex:
            with pytest.raises(ValueError) as gex:
                ol.run_window(gpu, params, reads, strand=strand, sisd=True)
            assert str(gex.value) == str(ex)
            continue
        _compare(gpu, orc, params, reads, strand=strand)


@pytest.mark.parametrize("params", [(ol.NW, 5, -4, -8, -6, -10, -4), (ol.SW, 5, -4, -8, -6, -10, -4),
                                    (ol.OV, 5, -4, -8, -6), (ol.NW, 3, -5, -3)])
def test_sisd_long_rows(gpu, orc, params):
    """Reads of 5-9 kb: rows of several workgroup tiles, and (global
    convex) graphs large enough for int32 rows."""
    rng = random.Random(hash(params) % 997)
    templ = "".join(rng.choice("ACGT") for _ in range(rng.randint(5000, 9000)))
    reads = [_mutate(rng, templ, 0.1, "ACGT") for _ in range(3)]
    _compare(gpu, orc, params, reads)


def test_sisd_overlap_end_cell_has_no_floor(gpu):
    """Where every overlap end cell is negative, the scalar engine reports
    the best one (the SIMD engine 0 and no alignment); values from the
    compiled reference."""
    for flag, want in ((False, (0, [])), (True, (-4, [(3, 0)]))):
        eng = gpu.engine(ol.OV, 5, -4, -8, -6, -10, -4, sisd=flag)
        gr = gpu.graph()
        try:
            p0, _, _ = gpu.align(eng, gr, "AAAA")
            gpu.add_alignment(gr, p0, "AAAA")
            pairs, score, width = gpu.align(eng, gr, "CC")
            assert (score, pairs) == want
            assert width == (32 if flag else 16)
        finally:
            gpu.free_graph(gr)
            gpu.free_engine(eng)


def test_sisd_overflow_bound_is_on_n_and_V(gpu):
    """A 16,777,200-base read on a one-node graph with linear gaps of -128:
    past the SIMD engine's bound on (n + 8, V), inside the scalar engine's
    on (n, V) -- the scalar engine aligns it; 8 bases more and it fails
    with its own message."""
    n = 16777200
    eng = gpu.engine(ol.NW, 1, -1, -128, sisd=True)
    simd = gpu.engine(ol.NW, 1, -1, -128)
    gr = gpu.graph()
    try:
        gpu.add_alignment(gr, [], "A")
        with pytest.raises(ValueError) as ex:
            gpu.align(simd, gr, "A" * n)
        assert str(ex.value) == "[spoa::SimdAlignmentEngine::Align] error: possible overflow!"
        pairs, score, width = gpu.align(eng, gr, "A" * n)
        assert score == 1 - 128 * (n - 1) and width == 32
        assert pairs[-1] == (0, n - 1) and len(pairs) == n
        assert all(p == (-1, c) for c, p in enumerate(pairs[:-1]))
        with pytest.raises(ValueError) as ex:
            gpu.align(eng, gr, "A" * (n + 8))
        assert str(ex.value) == "[spoa::SisdAlignmentEngine::Align] error: possible overflow!"
    finally:
        gpu.free_graph(gr)
        gpu.free_engine(eng)
        gpu.free_engine(simd)


@pytest.mark.parametrize("mode", ["batched", "resident"])
@pytest.mark.parametrize("params,strand", [((ol.NW, 5, -4, -8, -6, -10, -4), True), ((ol.OV, 5, -4, -8, -6), False),
                                           ((ol.SW, 5, -4, -8, -6, -10, -4), False), ((ol.NW, 3, -5, -3), False)])
def test_sisd_batched_and_resident_forms(orc, mode, params, strand):
    """Many windows at once (host-merged and device-resident graphs) with
    the scalar engine's alignments: the same consensus, MSA and kept
    strands as the oracle's scalar engine window by window."""
    rng = random.Random(41 + hash(params) % 1000)
    windows = []
    for _ in range(10):
        templ = "".join(rng.choice("ACGT") for _ in range(rng.randint(80, 400)))
        reads = [_mutate(rng, templ, 0.15, "ACGT") for _ in range(rng.randint(3, 9))]
        if strand:
            reads = [r if k % 2 == 0 else ol.revcomp(r) for k, r in enumerate(reads)]
        windows.append(reads)
    res = spoa_amd.run_windows(windows, *params, mode=mode, strand=strand, msa=True, sisd=True)
    for w, reads in enumerate(windows):
        cons, msa, recs = ol.run_window(orc, params, reads, strand=strand, sisd=True)
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w
        if strand:
            assert res.reversed[w] == [0 if r["kept_forward"] else 1 for r in recs], w


def test_sisd_selected_by_environment(orc, monkeypatch):
    """SPOA_HIP_ENGINE=sisd makes AlignmentEngine::Create (spoa_engine_new)
    stand in for the scalar engine, as a reference built without SIMD
    does; unset, the SIMD engine."""
    prod = common.product()
    for env, sisd in (("sisd", True), (None, False)):
        if env:
            monkeypatch.setenv("SPOA_HIP_ENGINE", env)
        else:
            monkeypatch.delenv("SPOA_HIP_ENGINE", raising=False)
        eng = prod.engine(ol.OV, 5, -4, -8, -6, -10, -4)
        gr = prod.graph()
        try:
            p0, _, _ = prod.align(eng, gr, "AAAA")
            prod.add_alignment(gr, p0, "AAAA")
            _, score, _ = prod.align(eng, gr, "CC")
            assert score == (-4 if sisd else 0)
        finally:
            prod.free_graph(gr)
            prod.free_engine(eng)
