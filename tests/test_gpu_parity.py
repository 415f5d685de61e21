"""GPU parity: the HIP alignment path against the oracle (a CPU
restatement of the reference, pinned to the compiled reference and to its
golden vectors by tests/test_oracle.py), bit-exact in every alignment pair,
score, consensus and MSA row."""
import random

import pytest

import common
import oracle_lib as ol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    return common.product()


@pytest.fixture(scope="module")
def orc():
    return ol.oracle()


def _compare(prod, orc, params, reads, weights=None, strand=False):
    cp, mp, rp = ol.run_window(prod, params, reads, weights, strand=strand)
    co, mo, ro = ol.run_window(orc, params, reads, weights, strand=strand)
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


@pytest.mark.parametrize("quality", [False, True])
@pytest.mark.parametrize("mode", range(9))
def test_reference_sample_modes(gpu, orc, mode, quality):
    """The reference's 18 end-to-end cases (test/spoa_test.cpp:154-440)."""
    params = common.MODES[mode]
    data = common.sample_reads()
    reads = [s for s, _ in data]
    weights = [[ord(c) - 33 for c in q] for _, q in data] if quality else None
    cons, msa = _compare(gpu, orc, params, reads, weights)
    golden = [g for g in common.golden_consensus().values()
              if g["params"] == list(params) and g["quality"] == quality]
    assert len(golden) == 1
    assert cons == golden[0]["consensus"]
    common.check_msa(msa, reads)


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


@pytest.mark.parametrize("seed", range(6))
def test_random_windows(gpu, orc, seed):
    """Seeded random windows over all 9 modes, random scores (including
    ones that force int32 scores), short and ragged reads, odd alphabets."""
    rng = random.Random(1000 + seed)
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
            ol.run_window(orc, params, reads, strand=strand)
        except ValueError as ex:
            with pytest.raises(ValueError) as gex:
                ol.run_window(gpu, params, reads, strand=strand)
            assert str(gex.value) == str(ex)
            continue
        _compare(gpu, orc, params, reads, strand=strand)


# (n, g, e, q, c) per subtype: gap penalties large enough that the
# reference's score bound picks int32 scores for graphs of a few hundred nodes
_WIDE_GAPS = {0: (-4, -60, -60, -60, -60), 1: (-4, -100, -40, -100, -40), 2: (-4, -100, -40, -120, -20)}


@pytest.mark.parametrize("sub", [0, 1, 2])
def test_local_int32_scores_on_int16_rows(gpu, orc, sub):
    """Local alignments that the reference scores at int32 although their
    scores fit int16 run on int16 rows with the int32 engine's row width
    (and, convex, its 8-column E/Q segments): the pairs, scores and reported
    width of the reference's int32 engine.  With a match score that leaves
    no int16 headroom (120 * (n + 8) > 30000) the jobs stay on int32 rows,
    also equal."""
    rng = random.Random(77 + sub)
    for m in (5, 120):
        params = (ol.SW, m) + _WIDE_GAPS[sub]
        templ = "".join(rng.choice("ACGT") for _ in range(300))
        reads = [_mutate(rng, templ, 0.12, "ACGT") for _ in range(6)]
        _, _, recs = ol.run_window(orc, params, reads)
        assert any(r["width"] == 32 for r in recs)
        _compare(gpu, orc, params, reads)
