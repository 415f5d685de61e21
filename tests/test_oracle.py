"""The oracle (oracle/spoa_oracle.cpp, a CPU restatement of the reference)
pinned against
  * the reference's own golden consensus strings and invariants
    (test/spoa_test.cpp:56-78, 154-440),
  * vectors written by the compiled reference (tests/golden/ref_vectors.json.gz),
  * the compiled reference itself where it is built (this container),
  * brute force on tiny graphs: the optimal score over every source-to-sink
    path with a textbook general-gap pairwise DP,
  * closed forms (the worst-case score bound and the width it selects)."""
import gzip
import hashlib
import itertools
import json
import os
import random

import pytest

import common
import oracle_lib as ol


@pytest.fixture(scope="module")
def orc():
    return ol.oracle()


@pytest.fixture(scope="module")
def vectors():
    with gzip.open(os.path.join(common.GOLDEN, "ref_vectors.json.gz"), "rt") as f:
        return json.load(f)


def pairs_hash(pairs):
    h = hashlib.sha1()
    for a, b in pairs:
        h.update(b"%d,%d;" % (a, b))
    return h.hexdigest()


def msa_hash(rows):
    return hashlib.sha1("\n".join(rows).encode()).hexdigest()


@pytest.mark.parametrize("name", sorted(common.golden_consensus()))
def test_reference_golden_consensus(orc, name):
    g = common.golden_consensus()[name]
    data = common.sample_reads()
    reads = [s for s, _ in data]
    weights = [[ord(c) - 33 for c in q] for _, q in data] if g["quality"] else None
    cons, msa, _ = ol.run_window(orc, g["params"], reads, weights)
    assert cons == g["consensus"]
    common.check_msa(msa, reads)


def _check_record(orc, rec, reads, weights=None, full=False):
    cons, msa, recs = ol.run_window(orc, rec["params"], reads, weights, strand=rec["strand"])
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
    assert cons == rec["consensus"]
    assert msa_hash(msa) == rec["msa_sha1"]
    if full:
        assert msa == rec["msa"]


def test_ref_vectors_sample(orc, vectors):
    data = common.sample_reads()
    reads = [s for s, _ in data]
    quals = [[ord(c) - 33 for c in q] for _, q in data]
    assert len(vectors["sample"]) == 18
    for rec in vectors["sample"]:
        _check_record(orc, rec, reads, quals if rec["quality"] else None)


@pytest.mark.parametrize("cfg", [1, 2, 3, 4])
def test_ref_vectors_configs(orc, vectors, cfg):
    recs = [r for r in vectors["configs"] if r["config"] == cfg]
    assert recs
    for rec in recs:
        _check_record(orc, rec, rec["reads"])


def test_ref_vectors_random(orc, vectors):
    for rec in vectors["random"]:
        if "error" in rec:
            with pytest.raises(ValueError) as ex:
                ol.run_window(orc, rec["params"], rec["reads"], strand=rec["strand"])
            assert str(ex.value) == rec["error"]
        else:
            _check_record(orc, rec, rec["reads"], full=True)


def _mutate(rng, s, p, alpha):
    out = []
    for b in s:
        if rng.random() < p:
            k = rng.randrange(3)
            if k == 0:
                out.append(rng.choice(alpha))
            elif k == 1:
                out.append(rng.choice(alpha))
                out.append(b)
        else:
            out.append(b)
    return "".join(out) or rng.choice(alpha)


def test_against_compiled_reference(orc):
    ref = ol.reference()
    if ref is None:
        pytest.skip("the compiled reference exists only where /root/reference was built")
    rng = random.Random(77)
    for _ in range(60):
        big = rng.random() < 0.4
        lo = -128 if big else -12
        params = (rng.randrange(3), rng.randint(0, 120 if big else 10), rng.randint(-120 if big else -10, 0),
                  rng.randint(lo, 0), rng.randint(lo, 0), rng.randint(lo, 0), rng.randint(lo, 0))
        alpha = rng.choice(["ACGT", "AC", "ACGTN"])
        tl = rng.choice([10, 60, 300])
        templ = "".join(rng.choice(alpha) for _ in range(tl))
        reads = [_mutate(rng, templ, 0.2, alpha) for _ in range(rng.randint(1, 5))]
        strand = rng.random() < 0.3
        try:
            a = ol.run_window(orc, params, reads, strand=strand)
        except ValueError as ex:
            with pytest.raises(ValueError) as ex2:
                ol.run_window(ref, params, reads, strand=strand)
            assert str(ex2.value) == str(ex)
            continue
        b = ol.run_window(ref, params, reads, strand=strand)
        assert a[0] == b[0] and a[1] == b[1]
        for x, y in zip(a[2], b[2]):
            assert x["pairs"] == y["pairs"] and x["score"] == y["score"]


# ------------------------------------------------------------ brute force --

def _gap(l, g, e, q, c):
    return 0 if l == 0 else max(g + (l - 1) * e, q + (l - 1) * c)


def _pairwise(path, read, typ, m, n, g, e, q, c):
    """Textbook DP with an arbitrary gap function (every gap length
    enumerated): NW global; SW local; OV as the reference scores it -- the
    graph prefix free, the read prefix gapped, the read suffix free."""
    P, S = len(path), len(read)
    NEG = -10 ** 9
    H = [[NEG] * (S + 1) for _ in range(P + 1)]
    H[0][0] = 0
    for i in range(1, P + 1):
        H[i][0] = 0 if typ in (ol.SW, ol.OV) else _gap(i, g, e, q, c)
    for j in range(1, S + 1):
        H[0][j] = 0 if typ == ol.SW else _gap(j, g, e, q, c)
    for i in range(1, P + 1):
        for j in range(1, S + 1):
            best = H[i - 1][j - 1] + (m if path[i - 1] == read[j - 1] else n)
            for k in range(1, i + 1):
                best = max(best, H[i - k][j] + _gap(k, g, e, q, c))
            for k in range(1, j + 1):
                best = max(best, H[i][j - k] + _gap(k, g, e, q, c))
            if typ == ol.SW:
                best = max(best, 0)
            H[i][j] = best
    if typ == ol.NW:
        return H[P][S]
    if typ == ol.OV:
        return max(0, max(H[P][1:]))
    return max(max(row) for row in H)


def _paths(chars, tails, heads):
    V = len(chars)
    outs = [[] for _ in range(V)]
    ins = [0] * V
    for t, h in zip(tails, heads):
        outs[t].append(h)
        ins[h] += 1
    res = []

    def walk(v, acc):
        acc = acc + chars[v]
        if not outs[v]:
            res.append(acc)
        for w in outs[v]:
            walk(w, acc)

    for v in range(V):
        if ins[v] == 0:
            walk(v, "")
    return res


def _structure(orc, gr):
    import ctypes
    V = orc.num_nodes(gr)
    chars = ctypes.create_string_buffer(max(1, V))
    cap = 4 * V + 16
    t = (ctypes.c_uint32 * cap)()
    h = (ctypes.c_uint32 * cap)()
    fn = orc.lib.oracle_graph_structure
    fn.restype = ctypes.c_int
    k = fn(ctypes.c_void_p(gr), chars, t, h, cap)
    return chars.raw[:V].decode(), list(t[:k]), list(h[:k])


@pytest.mark.parametrize("seed", range(4))
def test_brute_force_scores(orc, seed):
    """The oracle's score equals the best score over all source-to-sink
    paths (SW: over their subpaths) of the pairwise DP, in every mode."""
    rng = random.Random(seed)
    checked = 0
    while checked < 40:
        typ = rng.randrange(3)
        g = rng.randint(-6, -1)
        e = rng.randint(-6, -1)
        q = rng.randint(-8, -1)
        c = rng.randint(-3, -1)
        params = (typ, rng.randint(1, 5), rng.randint(-5, -1), g, e, q, c)
        eng = orc.engine(*params)
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
            read = _mutate(rng, base, 0.3, "ACG")
            _, score, _ = orc.align(eng, gr, read)
            # the engine normalises the scores as AlignmentEngine::Create does
            t, m, n, g, e, q, c = params
            if g >= e:
                e = g
                q, c = g, e
            elif g <= q or e >= c:
                q, c = g, e
            best = max(_pairwise(p, read, t, m, n, g, e, q, c) for p in paths)
            if t == ol.SW and best == 0:
                assert score is None
            else:
                assert score == best, (params, paths, read)
            checked += 1
        finally:
            orc.free_graph(gr)
            orc.free_engine(eng)


# ------------------------------------------------------------ closed forms --

def _worst(m, g, e, q, c, i, j):
    gap = lambda l: 0 if l == 0 else min(g + (l - 1) * e, q + (l - 1) * c)  # noqa: E731
    return min(-(m * min(i, j) + gap(abs(i - j))), gap(i) + gap(j))


@pytest.mark.parametrize("params,n,V,bits", [
    ((ol.NW, 3, -5, -3, -3, -3, -3), 500, 5000, 16),
    ((ol.NW, 5, -4, -8, -6, -10, -4), 10000, 20000, 32),
    ((ol.SW, 5, -4, -8, -6, -10, -4), 1500, 2000, 16),
    ((ol.SW, 5, -4, -8, -6, -10, -4), 2000, 5000, 32),
])
def test_width_choice(orc, params, n, V, bits):
    """int16 iff the worst-case score of (n + 8, V) stays above
    INT16_MIN + 1024 (src/alignment_engine.cpp:101-110, simd...:719-774)."""
    t, m, mm, g, e, q, c = params
    if g >= e:
        e = g
        q, c = g, e
    elif g <= q or e >= c:
        q, c = g, e
    w = _worst(m, g, e, q, c, n + 8, V)
    assert (w >= -32768 + 1024) == (bits == 16)
    # the oracle reaches the same decision on a chain graph of V nodes
    eng = orc.engine(*params)
    gr = orc.graph()
    try:
        orc.add_alignment(gr, [], "A" * V)
        _, _, width = orc.align(eng, gr, "A" * n)
        assert width == bits
    finally:
        orc.free_graph(gr)
        orc.free_engine(eng)
