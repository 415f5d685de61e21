"""Host-side graph code of the product on the CPU: spoa::Graph (the
drop-in) and the flat graph of the batched form, fed the oracle's
alignments, give the oracle's rank order, consensus and MSA (with and
without the consensus row), with unit and with quality weights; and both
report the reference's errors."""
import ctypes
import random

import pytest

import common
import oracle_lib as ol
import spoa_amd


def _flat(lib):
    f = {}
    for name in ("new", "free", "num_nodes", "add_alignment", "consensus", "msa", "ranks"):
        f[name] = getattr(lib, "spoa_poa_graph_" + name)
    f["new"].restype = ctypes.c_void_p
    f["free"].argtypes = [ctypes.c_void_p]
    f["num_nodes"].restype = ctypes.c_uint32
    f["num_nodes"].argtypes = [ctypes.c_void_p]
    f["add_alignment"].argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32), ctypes.c_uint32,
                                   ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                   ctypes.c_uint32]
    f["consensus"].argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32]
    f["msa"].argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64,
                         ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    f["ranks"].argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]
    return f


class FlatGraph:
    def __init__(self):
        self.f = _flat(spoa_amd.lib())
        self.h = self.f["new"]()

    def close(self):
        self.f["free"](self.h)

    def add(self, pairs, seq, weights=None):
        b = seq.encode()
        flat = (ctypes.c_int32 * max(1, 2 * len(pairs)))(*[x for p in pairs for x in p])
        w = None if weights is None else (ctypes.c_uint32 * len(weights))(*weights)
        rc = self.f["add_alignment"](self.h, flat, len(pairs), b, len(b), w, 0 if weights is None else len(weights))
        if rc != 0:
            raise ValueError(spoa_amd.lib().spoa_last_error().decode())

    def consensus(self):
        cap = 4 * self.f["num_nodes"](self.h) + 16
        buf = ctypes.create_string_buffer(cap)
        n = self.f["consensus"](self.h, buf, cap)
        return buf.raw[:n].decode()

    def msa(self, include_consensus=False):
        rows, cols = ctypes.c_uint32(0), ctypes.c_uint32(0)
        cap = 1 << 16
        while True:
            buf = ctypes.create_string_buffer(cap)
            if self.f["msa"](self.h, int(include_consensus), buf, cap, ctypes.byref(rows), ctypes.byref(cols)) == -2:
                cap = rows.value * cols.value + 1
                continue
            break
        c = cols.value
        return [buf.raw[k * c:(k + 1) * c].decode() for k in range(rows.value)]

    def ranks(self):
        v = self.f["num_nodes"](self.h)
        buf = (ctypes.c_uint32 * max(1, v))()
        n = self.f["ranks"](self.h, buf, v)
        return list(buf[:n])


def _window(orc, params, reads, weights):
    """Oracle alignments of every read, each against the graph of the
    reads before it."""
    eng = orc.engine(*params)
    gr = orc.graph()
    als = []
    try:
        for k, r in enumerate(reads):
            pairs, _, _ = orc.align(eng, gr, r)
            als.append(pairs)
            orc.add_alignment(gr, pairs, r, None if weights is None else weights[k])
        return als, orc.consensus(gr), orc.msa(gr), orc.ranks(gr)
    finally:
        orc.free_graph(gr)
        orc.free_engine(eng)


@pytest.mark.parametrize("mode", [0, 4, 8])
@pytest.mark.parametrize("quality", [False, True])
def test_graphs_match_oracle(mode, quality):
    orc = ol.oracle()
    prod = common.product()
    data = common.sample_reads()[:30]
    reads = [s for s, _ in data]
    weights = [[ord(c) - 33 for c in q] for _, q in data] if quality else None
    als, cons, msa, ranks = _window(orc, common.MODES[mode], reads, weights)
    # the drop-in spoa::Graph
    g = prod.graph()
    fg = FlatGraph()
    try:
        for k, (a, r) in enumerate(zip(als, reads)):
            prod.add_alignment(g, a, r, None if weights is None else weights[k])
            fg.add(a, r, None if weights is None else weights[k])
            # the rank order after every read (it fixes the DP row order)
            assert fg.ranks() == prod.ranks(g)
        assert prod.ranks(g) == ranks
        assert prod.consensus(g) == cons
        assert prod.msa(g) == msa
        assert fg.consensus() == cons
        assert fg.msa() == msa
        assert fg.msa(True)[:-1] == msa
        assert fg.msa(True)[-1].replace("-", "") == cons
        assert fg.msa(True) == prod.msa(g, True)
    finally:
        prod.free_graph(g)
        fg.close()


def test_graphs_random_windows():
    orc = ol.oracle()
    prod = common.product()
    rng = random.Random(5)
    for _ in range(20):
        params = (rng.randrange(3), rng.randint(1, 6), rng.randint(-6, -1), rng.randint(-8, -1),
                  rng.randint(-8, -1), rng.randint(-10, -1), rng.randint(-4, -1))
        alpha = rng.choice(["ACGT", "AC", "ACGTN"])
        templ = "".join(rng.choice(alpha) for _ in range(rng.randint(5, 80)))
        reads = []
        for _ in range(rng.randint(1, 8)):
            r = "".join(c for c in templ if rng.random() > 0.1)
            reads.append(r or "A")
        als, cons, msa, ranks = _window(orc, params, reads, None)
        g = prod.graph()
        fg = FlatGraph()
        try:
            for a, r in zip(als, reads):
                prod.add_alignment(g, a, r)
                fg.add(a, r)
            assert prod.ranks(g) == ranks == fg.ranks()
            assert prod.consensus(g) == cons == fg.consensus()
            assert prod.msa(g) == msa == fg.msa()
        finally:
            prod.free_graph(g)
            fg.close()


def test_graph_errors():
    """src/graph.cpp:163-194: the same messages from both graphs."""
    prod = common.product()
    g = prod.graph()
    fg = FlatGraph()
    try:
        prod.add_alignment(g, [], "ACGT")
        fg.add([], "ACGT")
        for bad, msg in (([(0, 7)], "invalid alignment"), ([(0, -1), (1, -1)], "missing sequence in alignment")):
            with pytest.raises(ValueError) as e1:
                prod.add_alignment(g, bad, "ACGT")
            with pytest.raises(ValueError) as e2:
                fg.add(bad, "ACGT")
            assert str(e1.value) == str(e2.value) == "[spoa::Graph::AddAlignment] error: " + msg
        with pytest.raises(ValueError) as e3:
            fg.add([(0, 0)], "ACGT", [1, 2])
        assert "unequal size" in str(e3.value)
        # an empty read adds nothing
        n = prod.num_nodes(g)
        prod.add_alignment(g, [], "")
        assert prod.num_nodes(g) == n
    finally:
        prod.free_graph(g)
        fg.close()


def test_create_errors():
    """AlignmentEngine::Create validation (src/alignment_engine.cpp:40-55)
    happens before any device is touched."""
    for args, msg in (((4, 1, -1, -1), "invalid alignment type!"),
                      ((0, 1, -1, 1), "gap opening penalty must be non-positive!"),
                      ((0, 1, -1, -1, 1), "gap extension penalty must be non-positive!")):
        with pytest.raises(ValueError) as ex:
            spoa_amd.AlignmentEngine.create(*args)
        assert str(ex.value).endswith(msg)
        assert str(ex.value).startswith("[spoa::AlignmentEngine::Create] error: ")
