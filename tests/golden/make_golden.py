"""Writes ref_vectors.json.gz: input/output vectors of the compiled
reference (oracle/_ref/libspoa_ref.so, built from /root/reference by
oracle/Makefile; SimdAlignmentEngine with AVX2 lane counts).

Run in the build container only.  The vectors pin the oracle in tests that
run anywhere, and through it the GPU path:
  sample  -- the reference's own test data (sample.fastq.gz) in its 9
             scoring modes, with and without quality weights;
  configs -- small seeded instances of the five benchmark window shapes;
  random  -- seeded random small windows over random scores, alphabets and
             strand modes.
Per alignment call: the score, the number of pairs and a SHA-1 of the
pairs (the pairs themselves for the random set); per window: consensus and
a SHA-1 of the MSA rows (the rows themselves for the random set)."""
import gzip
import hashlib
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

import common  # noqa: E402
import oracle_lib as ol  # noqa: E402


def pairs_hash(pairs):
    h = hashlib.sha1()
    for a, b in pairs:
        h.update(b"%d,%d;" % (a, b))
    return h.hexdigest()


def msa_hash(rows):
    return hashlib.sha1("\n".join(rows).encode()).hexdigest()


def record(ref, params, reads, weights=None, strand=False, full=False):
    cons, msa, recs = ol.run_window(ref, params, reads, weights, strand=strand)
    calls = []
    for r in recs:
        c = {"score": r["score"], "n": len(r["pairs"]), "sha1": pairs_hash(r["pairs"])}
        if full:
            c["pairs"] = r["pairs"]
        if strand:
            c["score_rc"] = r["score_rc"]
            c["sha1_rc"] = pairs_hash(r["pairs_rc"])
            c["kept_forward"] = r["kept_forward"]
        calls.append(c)
    out = {"params": list(params), "strand": strand, "calls": calls, "consensus": cons,
           "msa_sha1": msa_hash(msa), "msa_rows": len(msa)}
    if full:
        out["msa"] = msa
    return out


def mutate(rng, s, p, alpha):
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


def main():
    ref = ol.reference()
    assert ref is not None, "build the reference first: make -C oracle ref"
    data = common.sample_reads()
    reads = [s for s, _ in data]
    quals = [[ord(c) - 33 for c in q] for _, q in data]
    out = {"sample": [], "configs": [], "random": []}
    for params in common.MODES:
        for q in (False, True):
            rec = record(ref, params, reads, quals if q else None)
            rec["quality"] = q
            out["sample"].append(rec)
            print("sample", params, q, flush=True)

    import spoa_amd
    shapes = [(1, 1, 501), (2, 1, 502), (3, 8, 503), (4, 1, 504)]
    for cfg, nw, seed in shapes:
        c = spoa_amd.CONFIGS[cfg]
        ws = spoa_amd.config_windows(cfg, seed, nw)
        for w in range(nw):
            rd = ws.reads(w)
            rec = record(ref, c["params"], rd, strand=c.get("strand", False))
            rec.update({"config": cfg, "seed": seed, "window": w, "reads": rd})
            out["configs"].append(rec)
            print("config", cfg, w, len(rd), flush=True)

    rng = random.Random(20240601)
    while len(out["random"]) < 60:
        if rng.random() < 0.3:
            params = (rng.randrange(3), rng.randint(0, 120), rng.randint(-120, 0), rng.randint(-128, 0),
                      rng.randint(-128, 0), rng.randint(-128, 0), rng.randint(-128, 0))
        else:
            params = (rng.randrange(3), rng.randint(0, 10), rng.randint(-10, 0), rng.randint(-12, 0),
                      rng.randint(-12, 0), rng.randint(-14, 0), rng.randint(-12, 0))
        alpha = rng.choice(["ACGT", "AC", "ACGTN", "A"])
        tl = rng.choice([1, 5, 30, 120, 400])
        templ = "".join(rng.choice(alpha) for _ in range(tl))
        rd = []
        for _ in range(rng.randint(1, 6)):
            a = rng.randrange(tl)
            b = rng.randint(a + 1, tl)
            src = templ[a:b] if rng.random() < 0.5 else templ
            rd.append(mutate(rng, src, rng.random() * 0.3, alpha))
        strand = rng.random() < 0.25
        try:
            rec = record(ref, params, rd, strand=strand, full=True)
        except ValueError as ex:
            rec = {"params": list(params), "strand": strand, "error": str(ex)}
        rec["reads"] = rd
        out["random"].append(rec)
    path = os.path.join(HERE, "ref_vectors.json.gz")
    with gzip.open(path, "wt") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    main()
