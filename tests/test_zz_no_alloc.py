"""The entry point that allocates nothing (spoa_hip_align_noalloc, device
inputs and outputs, a fixed scratch array per GPU) against the plain
per-call entry point spoa_hip_align on the same rank views and reads: two
calls in a row, the second on a smaller input, and the oracle's alignment
for both.  The test allocates the inputs and outputs in device memory
before the calls, through the HIP runtime the library itself loads (torch
ships a HIP runtime of its own, which cannot share a process with it)."""
import ctypes
import random

import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


class HipGraph(ctypes.Structure):
    _fields_ = [("num_nodes", ctypes.c_uint32), ("row_char", ctypes.c_void_p), ("node_id", ctypes.c_void_p),
                ("pred_off", ctypes.c_void_p), ("pred_rank", ctypes.c_void_p), ("is_sink", ctypes.c_void_p)]


class HipParams(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int32), ("subtype", ctypes.c_int32), ("m", ctypes.c_int8),
                ("n", ctypes.c_int8), ("g", ctypes.c_int8), ("e", ctypes.c_int8), ("q", ctypes.c_int8),
                ("c", ctypes.c_int8)]


def _lib():
    import spoa_amd
    L = spoa_amd.lib()
    L.spoa_hip_engine_create.argtypes = [ctypes.c_int, ctypes.POINTER(HipParams), ctypes.POINTER(ctypes.c_void_p)]
    L.spoa_hip_engine_destroy.argtypes = [ctypes.c_void_p]
    L.spoa_hip_align.argtypes = [ctypes.c_void_p, ctypes.POINTER(HipGraph), ctypes.c_char_p, ctypes.c_uint32,
                                 ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                 ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.spoa_hip_align_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(HipGraph), ctypes.c_void_p, ctypes.c_uint32,
                                         ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    return L


def _window(seed, n_reads, length):
    rng = random.Random(seed)
    templ = "".join(rng.choice("ACGT") for _ in range(length))
    reads = []
    for _ in range(n_reads):
        out = []
        for b in templ:
            x = rng.random()
            if x < 0.04:
                out.append(rng.choice("ACGT"))
            elif x < 0.08:
                out.append(b + rng.choice("ACGT"))
            elif x >= 0.12:
                out.append(b)
        reads.append("".join(out))
    return reads


def _rank_view(orc, gr):
    """The rank-ordered CSR view of an oracle graph: rank_to_node order,
    predecessors in in-edge (creation) order, sink flags."""
    f = orc.lib.oracle_graph_structure
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32),
                  ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32]
    V = orc.num_nodes(gr)
    cap = 8 * V + 16
    chars = ctypes.create_string_buffer(max(1, V))
    tails = (ctypes.c_uint32 * cap)()
    heads = (ctypes.c_uint32 * cap)()
    ne = f(gr, chars, tails, heads, cap)
    assert ne >= 0
    ranks = orc.ranks(gr)
    rank_of = {v: r for r, v in enumerate(ranks)}
    inl = [[] for _ in range(V)]
    has_out = [False] * V
    for k in range(ne):
        inl[heads[k]].append(tails[k])
        has_out[tails[k]] = True
    row_char, pred_off, pred_rank, sink = [], [0], [], []
    for v in ranks:
        row_char.append(chars.raw[v])
        pred_rank.extend(rank_of[t] for t in inl[v])
        pred_off.append(len(pred_rank))
        sink.append(0 if has_out[v] else 1)
    return dict(V=V, row_char=row_char, node_id=list(ranks), pred_off=pred_off, pred_rank=pred_rank, sink=sink)


def _host_graph(v):
    keep = dict(row_char=(ctypes.c_uint8 * max(1, v["V"]))(*v["row_char"]),
                node_id=(ctypes.c_int32 * max(1, v["V"]))(*v["node_id"]),
                pred_off=(ctypes.c_uint32 * len(v["pred_off"]))(*v["pred_off"]),
                pred_rank=(ctypes.c_uint32 * max(1, len(v["pred_rank"])))(*v["pred_rank"]),
                sink=(ctypes.c_uint8 * max(1, v["V"]))(*v["sink"]))
    g = HipGraph(v["V"], *(ctypes.cast(keep[k], ctypes.c_void_p) for k in
                           ("row_char", "node_id", "pred_off", "pred_rank", "sink")))
    return g, keep


class Hip:
    """hipMalloc / hipMemcpy / hipFree / streams of the library's HIP runtime."""

    def __init__(self):
        import spoa_amd
        spoa_amd.lib()  # loads libamdhip64 first
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.rt.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        self.rt.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        self.bufs = []

    def put(self, ctype, values, n=None):
        n = max(1, len(values) if n is None else n)
        host = (ctype * n)(*values)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), ctypes.sizeof(host)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, host, ctypes.sizeof(host), 1) == 0  # host to device
        return p.value

    def get(self, ctype, ptr, n):
        host = (ctype * n)()
        assert self.rt.hipMemcpy(host, ctypes.c_void_p(ptr), ctypes.sizeof(host), 2) == 0  # device to host
        return list(host)

    def sync(self):
        assert self.rt.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _device_graph(hip, v):
    ptrs = [hip.put(ctypes.c_uint8, v["row_char"]), hip.put(ctypes.c_int32, v["node_id"]),
            hip.put(ctypes.c_uint32, v["pred_off"]), hip.put(ctypes.c_uint32, v["pred_rank"]),
            hip.put(ctypes.c_uint8, v["sink"])]
    return HipGraph(v["V"], *ptrs)


def _align_both(L, hip, eng, v, read, stream):
    # the plain entry point, host buffers
    hg, hkeep = _host_graph(v)
    cap = v["V"] + len(read)
    hp = (ctypes.c_int32 * (2 * cap))()
    np_, score, has = ctypes.c_uint32(0), ctypes.c_int32(0), ctypes.c_int32(0)
    rc = L.spoa_hip_align(eng, ctypes.byref(hg), read.encode(), len(read), hp, cap, ctypes.byref(np_),
                          ctypes.byref(score), ctypes.byref(has))
    assert rc == 0, L.spoa_hip_last_error()
    plain = ([(hp[2 * k], hp[2 * k + 1]) for k in range(np_.value)], score.value if has.value else None)
    # the no-allocation entry point, device buffers
    dg = _device_graph(hip, v)
    seq = hip.put(ctypes.c_uint8, list(read.encode()))
    pairs = hip.put(ctypes.c_int32, [-5] * (2 * cap))
    result = hip.put(ctypes.c_int32, [-9] * 5)
    hip.sync()
    rc = L.spoa_hip_align_noalloc(eng, ctypes.byref(dg), seq, len(read), pairs, cap, result, stream)
    assert rc == 0, L.spoa_hip_last_error()
    hip.sync()
    res = hip.get(ctypes.c_int32, result, 5)
    n = res[0]
    assert res[3] == 0
    p = hip.get(ctypes.c_int32, pairs, 2 * cap)
    got = ([(p[2 * k], p[2 * k + 1]) for k in range(n)], res[1] if res[2] else None)
    return plain, got, res[4]


# (Create arguments, the normalised spoa_hip_params: type, subtype, m, n, g, e, q, c)
CASES = [((ol.SW, 5, -4, -8, -6, -10, -4), (0, 2, 5, -4, -8, -6, -10, -4)),
         ((ol.NW, 3, -5, -3), (1, 0, 3, -5, -3, -3, -3, -3)),
         ((ol.OV, 5, -4, -8, -6), (2, 1, 5, -4, -8, -6, -8, -6))]


@pytest.mark.parametrize("params,hip_params", CASES)
def test_no_alloc_matches_plain_entry_point(params, hip_params):
    L = _lib()
    hip = Hip()
    orc = ol.oracle()
    oe = orc.engine(*params)
    p = HipParams(*hip_params)
    eng = ctypes.c_void_p()
    assert L.spoa_hip_engine_create(0, ctypes.byref(p), ctypes.byref(eng)) == 0
    st = ctypes.c_void_p()
    assert hip.rt.hipStreamCreate(ctypes.byref(st)) == 0
    try:
        # first call: a larger graph and read, on the engine's stream;
        # second: smaller ones, on a stream of the caller's
        for case, (seed, nreads, length, stream) in enumerate([(1, 12, 900, None), (2, 4, 300, st)]):
            reads = _window(seed, nreads, length)
            og = orc.graph()
            for r in reads[:-1]:
                pairs, _, _ = orc.align(oe, og, r)
                orc.add_alignment(og, pairs, r)
            v = _rank_view(orc, og)
            want = orc.align(oe, og, reads[-1])
            orc.free_graph(og)
            plain, got, width = _align_both(L, hip, eng, v, reads[-1], stream)
            assert plain == (want[0], want[1]), case
            assert got == plain, case
            assert width == want[2], case
    finally:
        hip.sync()
        hip.rt.hipStreamDestroy(st)
        hip.free()
        L.spoa_hip_engine_destroy(eng)
        orc.free_engine(oe)


def test_no_alloc_rejects_beyond_its_scratch():
    """A job whose matrices exceed the 64 MiB scratch fails before any
    launch with a message naming the limit."""
    L = _lib()
    hip = Hip()
    p = HipParams(1, 2, 5, -4, -8, -6, -10, -4)
    eng = ctypes.c_void_p()
    assert L.spoa_hip_engine_create(0, ctypes.byref(p), ctypes.byref(eng)) == 0
    try:
        V, n = 4000, 3000  # 4,001 rows x 5 matrices x 3,008 columns, at least 2 bytes each
        v = dict(V=V, row_char=[ord("A")] * V, node_id=list(range(V)),
                 pred_off=[0] + list(range(V)), pred_rank=list(range(V - 1)), sink=[0] * (V - 1) + [1])
        dg = _device_graph(hip, v)
        seq = hip.put(ctypes.c_uint8, [ord("A")] * n)
        pairs = hip.put(ctypes.c_int32, [0] * (2 * (V + n)))
        result = hip.put(ctypes.c_int32, [0] * 5)
        rc = L.spoa_hip_align_noalloc(eng, ctypes.byref(dg), seq, n, pairs, V + n, result, None)
        assert rc == 4  # SPOA_HIP_NO_MEMORY
        assert "64 MiB" in L.spoa_hip_last_error().decode()
    finally:
        hip.sync()
        hip.free()
        L.spoa_hip_engine_destroy(eng)
