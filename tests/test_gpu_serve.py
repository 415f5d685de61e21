"""The serving loop: jobs of two scoring parameter sets arrive over time,
are admitted into continuous batches and finish in any order; every job's
consensus and MSA are the same as when the job runs alone (the batched form
on that window alone, and the oracle)."""
import pytest

import oracle_lib as ol
import spoa_amd

pytestmark = pytest.mark.gpu


def _alone(reads, params):
    res = spoa_amd.run_windows(spoa_amd.Windows.from_lists([reads]), *params, mode="batched", msa=True)
    return res.consensus[0], res.msa[0]


def _check(jobs, results, oracle_every=1):
    for i, ((reads, params), r) in enumerate(zip(jobs, results)):
        assert r is not None and r.status == 0, (i, r and r.error)
        cons, msa = _alone(reads, params)
        assert r.consensus == cons, i
        assert r.msa == msa, i
        if i % oracle_every == 0:
            oc, om, _ = ol.run_window(ol.oracle(), params, reads)
            assert r.consensus == oc, i
            assert r.msa == om, i


def test_serve_mix_equals_alone():
    jobs, arrivals = spoa_amd.serving_jobs(7, 24, rate_per_s=60.0, template=(200, 900), reads=(5, 14))
    results, wall, stats = spoa_amd.serve_trace(jobs, arrivals, threads=8)
    assert stats["completed"] == len(jobs) and stats["failed"] == 0
    assert stats["max_active_seen"] > 1  # jobs overlapped
    _check(jobs, results, oracle_every=3)
    for r in results:
        assert r.submit_ms <= r.admit_ms <= r.done_ms


def test_serve_memory_admission_defers_and_stays_exact():
    """A tiny DP-memory budget: alignments wait for memory (deferred rounds),
    results are unchanged."""
    jobs, arrivals = spoa_amd.serving_jobs(11, 10, rate_per_s=1000.0, template=(300, 700), reads=(6, 10))
    results, wall, stats = spoa_amd.serve_trace(jobs, arrivals, threads=8, memory_fraction=2e-4)
    assert stats["completed"] == len(jobs)
    assert stats["deferred"] > 0
    _check(jobs, results, oracle_every=4)


def test_serve_rejects_bad_parameters():
    srv = spoa_amd.Server(threads=2)
    try:
        with pytest.raises(ValueError) as ex:
            srv.submit(["ACGT"], (spoa_amd.NW, 5, -4, 3))
        assert "gap opening penalty must be non-positive" in str(ex.value)
        jid = srv.submit([], (spoa_amd.NW, 5, -4, -8))
        r = srv.wait(jid, 30.0)
        assert r.status == 0 and r.consensus == ""
    finally:
        srv.close()


def _job(seed, template, n_reads, err=0.10):
    ws = spoa_amd.synth_windows(seed, 1, template=(template, template), reads=(n_reads, n_reads),
                                err=(err, err))
    return ws.reads(0)


def test_serve_size_classes_at_workload_sizes():
    """Jobs of the serving workload's sizes in every size-class lane: first
    reads of ~400 bp, ~1.2 kb, 2-5 kb and >= 5 kb, both scorings.  Jobs up
    to 6 kb equal the oracle run alone; a full-length job (10 kb template,
    30 reads, int32 scores, a graph of tens of thousands of nodes) equals
    the batched form run alone while the others share the device."""
    sw = (spoa_amd.SW, 5, -4, -8, -6, -10, -4)
    ov = (spoa_amd.OV, 5, -4, -8, -6)
    jobs = [
        (_job(1, 400, 8), sw), (_job(2, 450, 10), ov),
        (_job(3, 1200, 10), ov), (_job(4, 1500, 8), sw),
        (_job(5, 3000, 12), sw), (_job(6, 2500, 15), ov),
        (_job(7, 6000, 8), ov),
        (_job(8, 10000, 30, err=0.08), sw),
    ]
    lanes = [0, 0, 1, 1, 2, 2, 3, 3]
    arrivals = [0.05 * i for i in range(len(jobs))]
    results, wall, stats = spoa_amd.serve_trace(jobs, arrivals, threads=8)
    assert stats["completed"] == len(jobs) and stats["failed"] == 0
    assert [r.lane for r in results] == lanes
    assert stats["max_active_seen"] > 1
    for i, ((reads, params), r) in enumerate(zip(jobs, results)):
        assert r.status == 0, (i, r.error)
        cons, msa = _alone(reads, params)
        assert r.consensus == cons, i
        assert r.msa == msa, i
        if len(reads[0]) < 8000:
            oc, om, _ = ol.run_window(ol.oracle(), params, reads)
            assert r.consensus == oc, i
            assert r.msa == om, i
