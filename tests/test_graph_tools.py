"""The graph's windowed-alignment and output helpers against the compiled
reference, on the CPU (they are host code): Graph::Subgraph and
UpdateAlignment (include/spoa/graph.hpp:173-180, src/graph.cpp:518-605),
PrintDot (:607-647), the coverage-summary GenerateConsensus (:377-431), the
command-line tool's GFA (src/main.cpp:119-199), and its FASTA / FASTQ
reading.  Both graphs are built from the same alignments (the reference
engine's), so every difference is one of these functions."""
import ctypes
import os
import random

import pytest

import common
import oracle_lib as ol

REF = ol.reference()
pytestmark = pytest.mark.skipif(REF is None or not os.path.exists(common.PRODUCT_SO),
                                reason="needs the compiled reference and the product library")

I32P = ctypes.POINTER(ctypes.c_int32)
U32P = ctypes.POINTER(ctypes.c_uint32)


def _bind(chk):
    p = chk.prefix
    L = chk.lib
    f = getattr(L, p + "graph_subgraph")
    f.restype = ctypes.c_void_p
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, I32P, ctypes.c_uint32]
    f = getattr(L, p + "graph_update_alignment")
    f.argtypes = [ctypes.c_void_p, I32P, ctypes.c_uint32, I32P, ctypes.c_uint32]
    f = getattr(L, p + "graph_print_dot")
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    f = getattr(L, p + "graph_consensus_summary")
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint32, U32P, ctypes.c_uint32, U32P]
    f = getattr(L, p + "graph_gfa")
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_int,
                  ctypes.c_char_p, ctypes.c_uint64]
    return L


def _call(chk, name):
    return getattr(chk.lib, chk.prefix + name)


def _graphs(params, reads, weights=None, strand=False):
    """The same window in a product graph and a reference graph, merged from
    the reference engine's alignments; returns (product graph, reference
    graph, reversed flags)."""
    prod = common.product()
    eng = REF.engine(*params)
    rg, pg = REF.graph(), prod.graph()
    rev = []
    for k, r in enumerate(reads):
        pairs, score, _ = REF.align(eng, rg, r)
        seq, w = r, (weights[k] if weights else None)
        if strand:
            rc = ol.revcomp(r)
            pr, sr, _ = REF.align(eng, rg, rc)
            if not ((score or 0) >= (sr or 0)):
                pairs, seq = pr, rc
                w = list(reversed(w)) if w else None
            rev.append(seq is not r)
        REF.add_alignment(rg, pairs, seq, w)
        prod.add_alignment(pg, pairs, seq, w)
    REF.free_engine(eng)
    return prod, pg, rg, rev


def _window(seed, n_reads=8, length=150):
    rng = random.Random(seed)
    templ = "".join(rng.choice("ACGT") for _ in range(length))
    reads = []
    for _ in range(n_reads):
        out = []
        for b in templ:
            x = rng.random()
            if x < 0.05:
                out.append(rng.choice("ACGT"))
            elif x < 0.10:
                out.append(b + rng.choice("ACGT"))
            elif x >= 0.15:
                out.append(b)
        reads.append("".join(out))
    return reads


def _subgraph(chk, g, begin, end):
    n = chk.num_nodes(g)
    m = (ctypes.c_int32 * max(1, n))()
    sub = _call(chk, "graph_subgraph")(g, begin, end, m, n)
    assert sub, chk.error()
    k = chk.num_nodes(sub)
    return sub, list(m[:k])


@pytest.mark.parametrize("mode", [0, 4, 8])
def test_subgraph_and_update_alignment(mode):
    params = common.MODES[mode]
    _bind(REF)
    prod, pg, rg, _ = _graphs(params, _window(mode))
    _bind(prod)
    n = REF.num_nodes(rg)
    assert prod.num_nodes(pg) == n
    rng = random.Random(mode)
    eng = REF.engine(*params)
    for _ in range(6):
        begin = rng.randrange(0, n // 2)
        end = rng.randrange(begin, n)
        ps, pm = _subgraph(prod, pg, begin, end)
        rs, rm = _subgraph(REF, rg, begin, end)
        assert pm == rm
        assert prod.ranks(ps) == REF.ranks(rs)
        assert prod.consensus(ps) == REF.consensus(rs)
        # a read aligned to the subgraph, its pairs mapped back
        read = _window(mode + 100, 1)[0][begin // 2: begin // 2 + 60]
        pairs, _, _ = REF.align(eng, rs, read)
        flat_p = (ctypes.c_int32 * max(1, 2 * len(pairs)))(*[v for pr in pairs for v in pr])
        flat_r = (ctypes.c_int32 * max(1, 2 * len(pairs)))(*[v for pr in pairs for v in pr])
        mp = (ctypes.c_int32 * max(1, len(pm)))(*pm)
        assert _call(prod, "graph_update_alignment")(pg, mp, len(pm), flat_p, len(pairs)) == 0
        assert _call(REF, "graph_update_alignment")(rg, mp, len(pm), flat_r, len(pairs)) == 0
        assert list(flat_p[:2 * len(pairs)]) == list(flat_r[:2 * len(pairs)])
        assert all(flat_p[2 * i] == -1 or flat_p[2 * i] == pm[pairs[i][0]] for i in range(len(pairs)))
        prod.free_graph(ps)
        REF.free_graph(rs)
    REF.free_engine(eng)
    prod.free_graph(pg)
    REF.free_graph(rg)


def test_subgraph_rejects_nodes_out_of_range():
    prod, pg, rg, _ = _graphs(common.MODES[3], _window(5, 3, 60))
    _bind(prod)
    n = prod.num_nodes(pg)
    m = (ctypes.c_int32 * n)()
    assert not _call(prod, "graph_subgraph")(pg, 0, n, m, n)
    assert "out of range" in prod.error()


@pytest.mark.parametrize("quality", [False, True])
def test_dot_summary_and_gfa(tmp_path, quality):
    data = common.sample_reads()[:20]
    reads = [s for s, _ in data]
    weights = [[ord(c) - 33 for c in q] for _, q in data] if quality else None
    params = common.MODES[5]
    _bind(REF)
    prod, pg, rg, _ = _graphs(params, reads, weights)
    _bind(prod)
    for verbose in (0, 1):
        outs = []
        for chk, g in ((prod, pg), (REF, rg)):
            cons = ctypes.create_string_buffer(1 << 16)
            summ = (ctypes.c_uint32 * (1 << 16))()
            ln = ctypes.c_uint32(0)
            k = _call(chk, "graph_consensus_summary")(g, verbose, cons, 1 << 16, summ, 1 << 16, ctypes.byref(ln))
            assert k >= 0
            outs.append((cons.raw[:k], list(summ[:ln.value])))
        assert outs[0] == outs[1], verbose
        assert len(outs[0][1]) == len(outs[0][0]) * (1 if not verbose else 5)  # 4 codes + gaps
    texts = []
    for chk, g in ((prod, pg), (REF, rg)):
        path = str(tmp_path / (chk.prefix + "g.dot"))
        _call(chk, "graph_print_dot")(g, path.encode())
        texts.append(open(path).read())
    assert texts[0] == texts[1] and texts[0].startswith("digraph 20 {")
    headers = b"".join(("read%d" % i).encode() + b"\0" for i in range(len(reads)))
    for inc in (0, 1):
        out = []
        for chk, g in ((prod, pg), (REF, rg)):
            buf = ctypes.create_string_buffer(1 << 22)
            k = _call(chk, "graph_gfa")(g, headers, None, 0, inc, buf, 1 << 22)
            assert k > 0
            out.append(buf.raw[:k])
        assert out[0] == out[1]


def test_gfa_of_a_strand_window_marks_reversed_paths():
    reads = _window(11, 6, 120)
    reads = [r if i % 2 == 0 else ol.revcomp(r) for i, r in enumerate(reads)]
    params = (ol.NW, 5, -4, -8, -6, -10, -4)
    _bind(REF)
    prod, pg, rg, rev = _graphs(params, reads, strand=True)
    _bind(prod)
    assert any(rev)
    flags = bytes(int(x) for x in rev)
    headers = b"".join(("r%d" % i).encode() + b"\0" for i in range(len(reads)))
    out = []
    for chk, g in ((prod, pg), (REF, rg)):
        buf = ctypes.create_string_buffer(1 << 20)
        k = _call(chk, "graph_gfa")(g, headers, flags, len(rev), 1, buf, 1 << 20)
        out.append(buf.raw[:k])
    assert out[0] == out[1]
    assert b"-," in out[0] or b"-\t*" in out[0]


def test_cli_tool_without_a_device(tmp_path):
    """The command-line tool answers --version and rejects an unsupported
    file extension with the reference's message before it needs a device
    (its alignments and outputs are checked on the GPU by
    tests/test_gpu_cli.py)."""
    tool = os.path.join(common.ROOT, "spoa_amd", "spoa")
    assert os.path.exists(tool)
    import subprocess
    out = subprocess.run([tool, "--version"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip().startswith("spoa_amd")
    bad = subprocess.run([tool, str(tmp_path / "x.txt")], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "unsupported format extension" in bad.stderr
