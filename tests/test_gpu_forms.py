"""GPU parity of the launch configurations and of the batched form.

* Every launch configuration (columns per lane, waves per workgroup, rows
  split into several tiles) gives the same alignments as the oracle.
* The batched form (one launch per round over many windows) gives, window
  by window, the consensus and MSA of the compiled reference's vectors and
  of the oracle, and the plain per-call form gives the same."""
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


@pytest.fixture(scope="module")
def vectors():
    with gzip.open(os.path.join(common.GOLDEN, "ref_vectors.json.gz"), "rt") as f:
        return json.load(f)


def msa_hash(rows):
    return hashlib.sha1("\n".join(rows).encode()).hexdigest()


def _oracle_window(params, reads, strand=False):
    return ol.run_window(ol.oracle(), params, reads, strand=strand)


# (columns per lane, waves); with rows of up to ~700 columns, k=8 and one
# wave splits a row into two tiles
CONFIGS = [(8, 1), (16, 1), (32, 1), (8, 2), (8, 3), (16, 2), (16, 5), (8, 8)]


@pytest.mark.parametrize("k,waves", CONFIGS)
@pytest.mark.parametrize("mode", [0, 4, 8])
def test_launch_configs(k, waves, mode):
    params = common.MODES[mode]
    data = common.sample_reads()[:24]
    reads = [s for s, _ in data]
    cons_o, msa_o, recs_o = _oracle_window(params, reads)
    eng = spoa_amd.AlignmentEngine.create(*params)
    lib = spoa_amd.lib()
    for width in (16, 32):
        kk = k
        if params[3:] != (-8, -8, -8, -8) and params[3:] != (-8, -6, -8, -6) and width == 16 and k == 8:
            kk = 16  # convex int16 lanes hold whole 16-column segments
        assert lib.spoa_hip_set_config(eng.device, width, 0, 1 << 30, kk, waves) == 0
    gr = spoa_amd.Graph()
    for r, rec in zip(reads, recs_o):
        a, s = eng.align(r, gr)
        assert s == rec["score"]
        assert a == rec["pairs"]
        gr.add_alignment(a, r)
    assert gr.generate_consensus() == cons_o
    assert gr.generate_multiple_sequence_alignment() == msa_o


def test_batched_matches_reference_vectors(vectors):
    """Config-shaped windows recorded from the compiled reference."""
    for cfg in (1, 2, 3):
        recs = [r for r in vectors["configs"] if r["config"] == cfg]
        ws = spoa_amd.Windows.from_lists([r["reads"] for r in recs])
        res = spoa_amd.run_windows(ws, *recs[0]["params"], mode="batched", msa=True)
        for w, r in enumerate(recs):
            assert res.consensus[w] == r["consensus"], (cfg, w)
            assert msa_hash(res.msa[w]) == r["msa_sha1"], (cfg, w)


def test_per_call_matches_reference_vectors(vectors):
    for cfg in (1, 2):
        recs = [r for r in vectors["configs"] if r["config"] == cfg]
        ws = spoa_amd.Windows.from_lists([r["reads"] for r in recs])
        res = spoa_amd.run_windows(ws, *recs[0]["params"], mode="per_call", msa=True)
        for w, r in enumerate(recs):
            assert res.consensus[w] == r["consensus"]
            assert msa_hash(res.msa[w]) == r["msa_sha1"]


def test_strand_window_matches_reference_vectors(vectors):
    """Long global convex window, both strands per read (int32 scores)."""
    recs = [r for r in vectors["configs"] if r["config"] == 4]
    ws = spoa_amd.Windows.from_lists([r["reads"] for r in recs])
    res = spoa_amd.run_windows(ws, *recs[0]["params"], mode="batched", strand=True, msa=True)
    for w, r in enumerate(recs):
        assert res.consensus[w] == r["consensus"]
        assert msa_hash(res.msa[w]) == r["msa_sha1"]
        assert res.reversed[w] == [0 if c["kept_forward"] else 1 for c in r["calls"]]


@pytest.mark.parametrize("cfg", [1, 3])
def test_batched_matches_oracle_many_windows(cfg):
    """Many windows of uneven read counts in one batch, each equal to the
    oracle run alone."""
    c = spoa_amd.CONFIGS[cfg]
    ws = spoa_amd.config_windows(cfg, seed=4242, n_windows=24)
    res = spoa_amd.run_windows(ws, *c["params"], mode="batched", msa=True)
    for w in range(ws.n_windows):
        cons, msa, _ = _oracle_window(c["params"], ws.reads(w))
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w


def test_batched_local_int32_windows_beside_int16_ones():
    """One batch of local convex windows whose later reads the reference
    scores at int32 (graphs past its bound) next to reads it scores at
    int16: the int32 ones run on int16 rows with 8-column E/Q segments, a
    launch of their own beside the int16 one.  Every window equals the
    oracle run alone, with the reference's widths."""
    params = (spoa_amd.SW, 5, -4, -40, -25, -50, -12)
    rng = random.Random(2024)
    windows = []
    for w in range(10):
        templ = "".join(rng.choice("ACGT") for _ in range(rng.choice([250, 400])))
        reads = []
        for _ in range(rng.randint(8, 16)):
            out = []
            for b in templ:
                x = rng.random()
                if x < 0.08:
                    out.append(rng.choice("ACGT"))
                elif x < 0.16:
                    out.append(b + rng.choice("ACGT"))
                elif x >= 0.22:
                    out.append(b)
            reads.append("".join(out))
        windows.append(reads)
    ws = spoa_amd.Windows.from_lists(windows)
    res = spoa_amd.run_windows(ws, *params, mode="batched", msa=True)
    widths = set()
    for w, reads in enumerate(windows):
        cons, msa, recs = _oracle_window(params, reads)
        widths.update(r["width"] for r in recs)
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w
    assert {16, 32} <= widths  # (0: a read into the empty graph)


def test_batched_large_passes_match_oracle():
    """Passes of hundreds of jobs: every window of a 600-window batch equals
    the oracle run alone."""
    params = (spoa_amd.SW, 5, -4, -8, -6, -10, -4)
    rng = random.Random(606)
    windows = []
    for _ in range(600):
        templ = "".join(rng.choice("ACGT") for _ in range(rng.randint(60, 160)))
        reads = []
        for _ in range(rng.randint(2, 4)):
            out = []
            for ch in templ:
                x = rng.random()
                if x < 0.05:
                    out.append(rng.choice("ACGT"))
                elif x < 0.10:
                    out.append(ch + rng.choice("ACGT"))
                elif x >= 0.14:
                    out.append(ch)
            reads.append("".join(out) or templ)
        windows.append(reads)
    ws = spoa_amd.Windows.from_lists(windows)
    res = spoa_amd.run_windows(ws, *params, mode="batched", msa=True)
    for w, reads in enumerate(windows):
        cons, msa, _ = _oracle_window(params, reads)
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w


def test_batched_full_size_batch_matches_oracle():
    """A full 2,048-window batch of 1.5 kb local convex windows (passes of
    1,024 jobs whose staging blocks pass 128 MiB in the later rounds, so
    they are packed and copied in in slices): windows spread over the
    batch equal the oracle run alone."""
    c = spoa_amd.CONFIGS[2]
    ws = spoa_amd.config_windows(2, seed=31, n_windows=2048)
    res = spoa_amd.run_windows(ws, *c["params"], mode="batched", msa=True)
    for w in (0, 700, 1401, 2047):
        cons, msa, _ = _oracle_window(c["params"], ws.reads(w))
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w


def test_context_reuse_and_budget_split():
    """A reused batched context gives identical results run after run, also
    when a small memory budget splits every round into several passes."""
    c = spoa_amd.CONFIGS[3]
    ws = spoa_amd.config_windows(3, seed=99, n_windows=12)
    ref = spoa_amd.run_windows(ws, *c["params"], mode="batched", msa=True)
    ctx = spoa_amd.BatchContext(*c["params"], msa=True)
    for _ in range(2):
        res = ctx.run(ws)
        assert res.consensus == ref.consensus
        assert res.msa == ref.msa
    # about three jobs' matrices per pass
    assert spoa_amd.lib().spoa_hip_set_memory_budget(ctx.device, 3 * 4 * 1024 * 1024 * 3) == 0
    res = ctx.run(ws)
    assert res.consensus == ref.consensus
    assert res.msa == ref.msa


def test_autotuner_picks_configs_and_keeps_results(tmp_path):
    """The autotuner times launch configurations per score width (the
    int32 width, which only late rounds reach, on the whole batch), installs
    the fastest, caches it, and the batched results do not change."""
    c = spoa_amd.CONFIGS[2]
    ws = spoa_amd.config_windows(2, seed=5, n_windows=16)
    ref = spoa_amd.run_windows(ws, *c["params"], mode="batched", msa=True)
    ctx = spoa_amd.BatchContext(*c["params"], msa=True)
    cache = str(tmp_path / "tune.json")
    tuned = ctx.autotune(ws, rounds=3, cache=cache)
    widths = {t["width"]: t for t in tuned}
    assert set(widths) == {16, 32}
    t16 = widths[16]
    assert t16["cols_per_lane"] in (16, 32) and 1 <= t16["waves"] <= 8  # convex int16 lanes hold 16+
    assert t16["kernel_ms"] > 0 and not t16["cached"]
    # the int32-width jobs (graphs past the int16 bound) exist only in the
    # late rounds: their launch is timed on the whole batch
    t32 = widths[32]
    assert t32["whole_batch"] and t32["kernel_ms"] > 0 and t32["cols_per_lane"] in (8, 16)
    res = ctx.run(ws)
    assert res.consensus == ref.consensus
    assert res.msa == ref.msa
    # a second context reads the cache instead of measuring
    ctx2 = spoa_amd.BatchContext(*c["params"], msa=True)
    tuned2 = ctx2.autotune(ws, rounds=3, cache=cache)
    t2 = {t["width"]: t for t in tuned2}[16]
    assert t2["cached"] and (t2["cols_per_lane"], t2["waves"]) == (t16["cols_per_lane"], t16["waves"])
    res2 = ctx2.run(ws)
    assert res2.consensus == ref.consensus


def test_graph_mode_replays_match_stream_mode():
    """Every pass launched as a captured HIP graph: the same results as the
    stream launches, run after run, with eager stream-mode runs in between.
    Passes of the same shape as the engine's previous one update its
    executable graph in place and replay it."""
    c = spoa_amd.CONFIGS[2]
    ws = spoa_amd.config_windows(2, seed=8, n_windows=12)
    eager = spoa_amd.BatchContext(*c["params"], msa=True)
    graph = spoa_amd.BatchContext(*c["params"], msa=True, graph=True)
    ref = eager.run(ws)
    for _ in range(3):
        res = graph.run(ws)
        assert res.consensus == ref.consensus
        assert res.msa == ref.msa
        assert res.stats["graph_updates"] > 0
        # a pass per half per round, but for the first round (every read
        # then meets an empty graph: nothing to launch)
        assert res.stats["graph_builds"] + res.stats["graph_updates"] == (res.stats["rounds"] - 1) * 2
        again = eager.run(ws)
        assert again.consensus == ref.consensus


def test_per_call_one_engine_per_host_thread():
    """The per-call drop-in on several host threads, an engine each (the
    reference's threading model): the same windows as the batched form."""
    c = spoa_amd.CONFIGS[3]
    ws = spoa_amd.config_windows(3, seed=12, n_windows=8)
    ref = spoa_amd.run_windows(ws, *c["params"], mode="batched", msa=True)
    res = spoa_amd.run_windows(ws, *c["params"], mode="per_call", msa=True, threads=4)
    assert res.consensus == ref.consensus
    assert res.msa == ref.msa

