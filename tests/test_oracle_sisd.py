"""The oracle's restatement of the reference's scalar engine
(SisdAlignmentEngine, src/sisd_alignment_engine.cpp), the engine that
AlignmentEngine::Create returns from a reference built without SIMD
(src/alignment_engine.cpp:68-71).

Pinned by: the reference's 18 golden consensus strings
(test/spoa_test.cpp:154-440, which the scalar engine passes too); vectors
recorded from the compiled scalar engine (tests/golden/ref_sisd_vectors.json.gz,
written by tests/golden/make_golden_sisd.py); the compiled scalar engine
itself where it was built; brute-force optimal scores on tiny graphs (its
overlap end cell has no 0 floor); and the closed-form overflow bound on
(n, V) instead of the SIMD engine's (n + 8, V)."""
import gzip
import json
import os
import random

import pytest

import common
import oracle_lib as ol
from test_oracle import _mutate, _pairwise, _paths, _structure, msa_hash, pairs_hash


@pytest.fixture(scope="module")
def orc():
    return ol.oracle()


@pytest.fixture(scope="module")
def vectors():
    with gzip.open(os.path.join(common.GOLDEN, "ref_sisd_vectors.json.gz"), "rt") as f:
        return json.load(f)


@pytest.mark.parametrize("name", sorted(common.golden_consensus()))
def test_sisd_reference_golden_consensus(orc, name):
    g = common.golden_consensus()[name]
    data = common.sample_reads()
    reads = [s for s, _ in data]
    weights = [[ord(c) - 33 for c in q] for _, q in data] if g["quality"] else None
    cons, msa, _ = ol.run_window(orc, g["params"], reads, weights, sisd=True)
    assert cons == g["consensus"]
    common.check_msa(msa, reads)


def _check(orc, rec, reads, weights=None, full=False):
    cons, msa, recs = ol.run_window(orc, rec["params"], reads, weights, strand=rec["strand"], sisd=True)
    assert len(recs) == len(rec["calls"])
    for k, (got, exp) in enumerate(zip(recs, rec["calls"])):
        assert got["score"] == exp["score"], k
        assert len(got["pairs"]) == exp["n"], k
        assert pairs_hash(got["pairs"]) == exp["sha1"], k
        if full:
            assert [list(p) for p in got["pairs"]] == exp["pairs"], k
        if rec["strand"]:
            assert got["score_rc"] == exp["score_rc"]
            assert pairs_hash(got["pairs_rc"]) == exp["sha1_rc"]
            assert got["kept_forward"] == exp["kept_forward"]
        assert got["width"] in (0, 32)  # int32 scores (0: no DP ran, an empty graph)
    assert cons == rec["consensus"]
    assert msa_hash(msa) == rec["msa_sha1"]
    if full:
        assert msa == rec["msa"]


def test_sisd_vectors_sample(orc, vectors):
    data = common.sample_reads()
    reads = [s for s, _ in data]
    quals = [[ord(c) - 33 for c in q] for _, q in data]
    assert len(vectors["sample"]) == 18
    for rec in vectors["sample"]:
        _check(orc, rec, reads, quals if rec["quality"] else None)


@pytest.mark.parametrize("cfg", [1, 2, 3, 4])
def test_sisd_vectors_configs(orc, vectors, cfg):
    recs = [r for r in vectors["configs"] if r["config"] == cfg]
    assert recs
    for rec in recs:
        _check(orc, rec, rec["reads"])


def test_sisd_vectors_random(orc, vectors):
    for rec in vectors["random"]:
        if "error" in rec:
            with pytest.raises(ValueError) as ex:
                ol.run_window(orc, rec["params"], rec["reads"], strand=rec["strand"], sisd=True)
            assert str(ex.value) == rec["error"]
        else:
            _check(orc, rec, rec["reads"], full=True)


def test_sisd_against_compiled_reference(orc):
    ref = ol.reference()
    if ref is None:
        pytest.skip("the compiled reference exists only where /root/reference was built")
    rng = random.Random(91)
    for _ in range(80):
        big = rng.random() < 0.3
        lo = -128 if big else -12
        params = (rng.randrange(3), rng.randint(0, 120 if big else 10), rng.randint(-120 if big else -10, 0),
                  rng.randint(lo, 0), rng.randint(lo, 0), rng.randint(lo, 0), rng.randint(lo, 0))
        alpha = rng.choice(["ACGT", "AC", "ACGTN"])
        tl = rng.choice([10, 60, 300])
        templ = "".join(rng.choice(alpha) for _ in range(tl))
        reads = [_mutate(rng, templ, 0.2, alpha) if rng.random() < 0.85 else _mutate(rng, templ[: tl // 3], 0.4, alpha)
                 for _ in range(rng.randint(1, 6))]
        strand = rng.random() < 0.3
        try:
            a = ol.run_window(orc, params, reads, strand=strand, sisd=True)
        except ValueError as ex:
            with pytest.raises(ValueError) as ex2:
                ol.run_window(ref, params, reads, strand=strand, sisd=True)
            assert str(ex2.value) == str(ex)
            continue
        b = ol.run_window(ref, params, reads, strand=strand, sisd=True)
        assert a[0] == b[0] and a[1] == b[1]
        for x, y in zip(a[2], b[2]):
            assert x["pairs"] == y["pairs"] and x["score"] == y["score"]


@pytest.mark.parametrize("seed", range(3))
def test_sisd_brute_force_scores(orc, seed):
    """The scalar engine's score is the best score over all source-to-sink
    paths of the pairwise DP; its overlap end cell has no 0 floor
    (sisd...:296-311), so negative overlap optima are reported as they are."""
    rng = random.Random(1000 + seed)
    checked = negative_ov = 0
    while checked < 40:
        typ = rng.randrange(3)
        g, e = rng.randint(-6, -1), rng.randint(-6, -1)
        q, c = rng.randint(-8, -1), rng.randint(-3, -1)
        params = (typ, rng.randint(1, 5), rng.randint(-5, -1), g, e, q, c)
        eng = orc.engine(*params, sisd=True)
        gr = orc.graph()
        try:
            base = "".join(rng.choice("ACG") for _ in range(rng.randint(2, 5)))
            for _ in range(rng.randint(1, 3)):
                r = _mutate(rng, base, 0.4, "ACG")
                pairs, _, _ = orc.align(eng, gr, r)
                orc.add_alignment(gr, pairs, r)
            chars, tails, heads = _structure(orc, gr)
            paths = _paths(chars, tails, heads)
            if len(paths) > 64:
                continue
            read = _mutate(rng, base, 0.3, "ACG") if rng.random() < 0.7 else "T" * rng.randint(1, 3)
            _, score, _ = orc.align(eng, gr, read)
            t, m, n, g, e, q, c = params
            if g >= e:
                e = g
                q, c = g, e
            elif g <= q or e >= c:
                q, c = g, e
            if t == ol.OV:
                best = max(max(_ov_row(p, read, m, n, g, e, q, c)) for p in paths)
                negative_ov += best < 0
            else:
                best = max(_pairwise(p, read, t, m, n, g, e, q, c) for p in paths)
            if t == ol.SW and best == 0:
                assert score is None
            else:
                assert score == best, (params, paths, read)
            checked += 1
        finally:
            orc.free_graph(gr)
            orc.free_engine(eng)
    assert checked == 40


def _ov_row(path, read, m, n, g, e, q, c):
    """The last row of the overlap DP over one path (columns 1..n), the
    cells the scalar engine's overlap end cell ranges over."""
    import test_oracle as to
    P, S = len(path), len(read)
    NEG = -10 ** 9
    H = [[NEG] * (S + 1) for _ in range(P + 1)]
    H[0][0] = 0
    for i in range(1, P + 1):
        H[i][0] = 0
    for j in range(1, S + 1):
        H[0][j] = to._gap(j, g, e, q, c)
    for i in range(1, P + 1):
        for j in range(1, S + 1):
            best = H[i - 1][j - 1] + (m if path[i - 1] == read[j - 1] else n)
            for k in range(1, i + 1):
                best = max(best, H[i - k][j] + to._gap(k, g, e, q, c))
            for k in range(1, j + 1):
                best = max(best, H[i][j - k] + to._gap(k, g, e, q, c))
            H[i][j] = best
    return H[P][1:]


@pytest.mark.parametrize("graph,read,params,simd,sisd", [
    ("AAAA", "CC", (ol.OV, 5, -4, -8, -6, -10, -4), (0, []), (-4, [(3, 0)])),
    ("AAAAAAAA", "C", (ol.OV, 5, -4, -8, -6), (0, []), (-4, [(7, 0)])),
    ("GGGG", "CCCCCC", (ol.OV, 3, -5, -3), (0, []), (-3, [(-1, 0)])),
    ("ACGT", "TT", (ol.OV, 5, -4, -8, -6, -10, -4), (5, [(3, 0)]), (5, [(3, 0)])),
])
def test_sisd_overlap_end_cell_has_no_floor(orc, graph, read, params, simd, sisd):
    """Where every overlap end cell is negative, the SIMD engine reports
    score 0 and an empty alignment (its row maxima are floored at 0,
    simd...:252-262, and no column holds 0), the scalar engine the best
    negative cell and its path (values from the compiled reference)."""
    for flag, want in ((False, simd), (True, sisd)):
        eng = orc.engine(*params, sisd=flag)
        gr = orc.graph()
        try:
            p0, _, _ = orc.align(eng, gr, graph)
            orc.add_alignment(gr, p0, graph)
            pairs, score, _ = orc.align(eng, gr, read)
            assert (score, pairs) == want
        finally:
            orc.free_graph(gr)
            orc.free_engine(eng)


def test_sisd_overflow_bound_is_on_n_and_V(orc):
    """The scalar engine rejects a read when WorstCaseAlignmentScore(n, V)
    passes INT32_MIN + 1024 (sisd...:269-272), the SIMD engine already at
    (n + 8, V): a one-node graph and linear gaps of -128 put the two bounds
    at n = 16,777,208 and n = 16,777,200."""
    params = (ol.NW, 1, -1, -128)
    for n, simd_ok, sisd_ok in ((16777199, True, True), (16777200, False, True), (16777208, False, False)):
        for flag, ok in ((False, simd_ok), (True, sisd_ok)):
            eng = orc.engine(*params, sisd=flag)
            gr = orc.graph()
            try:
                orc.add_alignment(gr, [], "A")
                if ok:
                    pairs, score, width = orc.align(eng, gr, "A" * n)
                    assert len(pairs) == n and score == 1 - 128 * (n - 1)
                    assert width == 32
                else:
                    with pytest.raises(ValueError) as ex:
                        orc.align(eng, gr, "A" * n)
                    who = "Sisd" if flag else "Simd"
                    assert str(ex.value) == "[spoa::%sAlignmentEngine::Align] error: possible overflow!" % who
            finally:
                orc.free_graph(gr)
                orc.free_engine(eng)
