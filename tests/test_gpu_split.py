"""GPU parity of split jobs: a row spanning several tiles, every tile on a
workgroup of its own, the tiles handing each row's boundary (the last H and
the gap carry) to the next through device memory.

* Per call: every split launch shape (columns per lane x waves per tile),
  every scoring mode family, against the oracle's pairs and scores.
* Batched and device-resident windows with split launches, against the
  oracle run alone on each window.
* Long global convex windows (int32, the default split into one-wave tiles)
  against the oracle, per call and strand-ambiguous in a batch.
* A wait limit of zero makes a split job fail with the time-out error (or
  finish correctly, when no tile ever had to wait), and the engine works
  again with the default limit."""
import random

import pytest

import common
import oracle_lib as ol
import spoa_amd

pytestmark = pytest.mark.gpu


def _reads(seed, template_len, n_reads, err=0.1):
    rng = random.Random(seed)
    templ = "".join(rng.choice("ACGT") for _ in range(template_len))
    reads = []
    for _ in range(n_reads):
        out = []
        for b in templ:
            x = rng.random()
            if x < err / 3:
                out.append(rng.choice("ACGT"))
            elif x < 2 * err / 3:
                out.append(b + rng.choice("ACGT"))
            elif x >= err:
                out.append(b)
        reads.append("".join(out))
    return reads


def _per_call_against_oracle(params, reads, configs):
    """Aligns reads into one graph through the drop-in engine with the given
    (width, k, waves, split) overrides, each call against the oracle."""
    cons_o, msa_o, recs_o = ol.run_window(ol.oracle(), params, reads)
    eng = spoa_amd.AlignmentEngine.create(*params)
    lib = spoa_amd.lib()
    for width, k, waves, split in configs:
        assert lib.spoa_hip_set_config_split(eng.device, width, 0, 1 << 30, k, waves, split) == 0, \
            spoa_amd.lib().spoa_hip_last_error()
    gr = spoa_amd.Graph()
    for i, (r, rec) in enumerate(zip(reads, recs_o)):
        a, s = eng.align(r, gr)
        assert s == rec["score"], i
        assert a == rec["pairs"], i
        gr.add_alignment(a, r)
    assert gr.generate_consensus() == cons_o
    assert gr.generate_multiple_sequence_alignment() == msa_o


# (columns per lane, waves per tile); reads of about 2.5 kb make rows of 2
# (32 x 1, 16 x 2) to 5 (8 x 1) tiles
SPLIT_SHAPES = [(8, 1), (16, 1), (8, 2), (16, 2), (32, 1)]


@pytest.mark.parametrize("k,waves", SPLIT_SHAPES)
@pytest.mark.parametrize("mode", [0, 2, 4, 5, 7, 8])
def test_per_call_split_shapes(k, waves, mode):
    params = common.MODES[mode]
    reads = _reads(100 + mode, 2500, 5, err=0.12)
    convex = params[3:] not in ((-8, -8, -8, -8), (-8, -6, -8, -6))
    cfgs = []
    for width in (16, 32):
        kk = k
        if convex and width == 16 and k == 8:
            kk = 16  # convex int16 lanes hold whole 16-column segments
        if width == 32 and kk == 32:
            kk = 16
        cfgs.append((width, kk, waves, 1))
    _per_call_against_oracle(params, reads, cfgs)


def test_per_call_split_sample_reads_all_modes():
    """The reference's sample reads in all 9 modes, every row split into
    one-wave tiles of 8 (int16 convex: 16) columns."""
    data = common.sample_reads()[:12]
    reads = [s for s, _ in data]
    for mode, params in enumerate(common.MODES):
        convex = params[3:] not in ((-8, -8, -8, -8), (-8, -6, -8, -6))
        cfgs = [(16, 16 if convex else 8, 1, 1), (32, 8, 1, 1)]
        _per_call_against_oracle(params, reads, cfgs)


def test_per_call_long_global_convex_default_split():
    """A 6 kb global convex window at int32 (rows of ~12 one-wave tiles by
    default) against the oracle."""
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    reads = _reads(7, 6000, 4, err=0.1)
    eng = spoa_amd.AlignmentEngine.create(*params)
    k, w, sp = (spoa_amd.ctypes.c_int32(), spoa_amd.ctypes.c_int32(), spoa_amd.ctypes.c_int32())
    assert spoa_amd.lib().spoa_hip_config_split(eng.device, 6000, 32, 0, spoa_amd.ctypes.byref(k),
                                                spoa_amd.ctypes.byref(w), spoa_amd.ctypes.byref(sp)) == 0
    assert sp.value == 1 and w.value == 1
    _per_call_against_oracle(params, reads, [])


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("cfg", [1, 3])
def test_batched_split_matches_oracle(cfg, resident):
    """Config-1 and config-3 windows with every row split into one-wave
    tiles of 8 columns (int16 rows of 400-600 columns: 1-2 tiles), batched
    and device-resident, each window equal to the oracle run alone."""
    c = spoa_amd.CONFIGS[cfg]
    ws = spoa_amd.config_windows(cfg, seed=777, n_windows=16)
    ctx = spoa_amd.BatchContext(*c["params"], msa=True, resident=resident)
    for width in (16, 32):
        ctx.set_config(width, 8, 1, split=True)
    res = ctx.run(ws)
    for w in range(ws.n_windows):
        cons, msa, _ = ol.run_window(ol.oracle(), c["params"], ws.reads(w))
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w


@pytest.mark.parametrize("resident", [False, True])
def test_batched_local_convex_split(resident):
    """Local convex 1.5 kb windows (config-2 scoring; reads of 1-2 kb) with
    rows split into one-wave tiles of 16 columns, both score widths."""
    c = spoa_amd.CONFIGS[2]
    ws = spoa_amd.synth_windows(31, 6, template=(1500, 1500), reads=(12, 12), read_len=(1000, 2000),
                                err=(0.15, 0.15))
    ctx = spoa_amd.BatchContext(*c["params"], msa=True, resident=resident)
    ctx.set_config(16, 16, 1, split=True)
    ctx.set_config(32, 8, 2, split=True)
    res = ctx.run(ws)
    for w in range(ws.n_windows):
        cons, msa, _ = ol.run_window(ol.oracle(), c["params"], ws.reads(w))
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w


@pytest.mark.parametrize("resident", [False, True])
def test_batched_strand_long_windows_default_split(resident):
    """Config-4-shaped windows (global convex, int32, half the reads
    reverse-complemented, strand-ambiguous) with the default split launch,
    each window equal to the oracle run alone."""
    c = spoa_amd.CONFIGS[4]
    ws = spoa_amd.synth_windows(4040, 3, template=(5000, 6000), reads=(5, 6), err=(0.1, 0.1), revcomp_odd=True)
    res = spoa_amd.run_windows(ws, *c["params"], mode="resident" if resident else "batched", strand=True,
                               msa=True)
    for w in range(ws.n_windows):
        cons, msa, recs = ol.run_window(ol.oracle(), c["params"], ws.reads(w), strand=True)
        assert res.consensus[w] == cons, w
        assert res.msa[w] == msa, w
        assert res.reversed[w] == [0 if r["kept_forward"] else 1 for r in recs], w


def test_wait_limit_zero_fails_cleanly_then_recovers():
    """With no time allowed for waiting, a split job either fails with the
    time-out error or (no tile ever waited) aligns correctly; with the
    default limit back the same engine aligns correctly."""
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    reads = _reads(9, 3000, 3, err=0.1)
    cons_o, msa_o, recs_o = ol.run_window(ol.oracle(), params, reads)
    eng = spoa_amd.AlignmentEngine.create(*params)
    L = spoa_amd.lib()
    assert L.spoa_hip_set_config_split(eng.device, 32, 0, 1 << 30, 8, 1, 1) == 0
    gr = spoa_amd.Graph()
    a, s = eng.align(reads[0], gr)
    gr.add_alignment(a, reads[0])
    assert L.spoa_hip_set_wait_limit(eng.device, 0) == 0
    try:
        a1, s1 = eng.align(reads[1], gr)
        assert (a1, s1) == (recs_o[1]["pairs"], recs_o[1]["score"])
    except Exception as ex:  # the time-out, reported by the engine
        assert "timed out waiting for each other" in str(ex)
    assert L.spoa_hip_set_wait_limit(eng.device, 1000000000) == 0
    a1, s1 = eng.align(reads[1], gr)
    assert (a1, s1) == (recs_o[1]["pairs"], recs_o[1]["score"])
