"""ctypes bindings to the checkers: the oracle (oracle/liboracle.so, a CPU
restatement of the reference) and, where it was built in this container,
the reference itself (oracle/_ref/libspoa_ref.so).

Test infrastructure only: the product never imports this module.
Both libraries export the same entry points under the prefixes ``oracle_``
and ``ref_``, so one driver runs either of them.
"""
import ctypes
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libspoa_ref.so")

SW, NW, OV = 0, 1, 2

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_u32p = ctypes.POINTER(ctypes.c_uint32)


class Checker:
    """One of the two checker libraries, driven through its C shim."""

    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        f = self._fn
        f("last_error").restype = ctypes.c_char_p
        f("engine_new").restype = ctypes.c_void_p
        f("engine_new").argtypes = [ctypes.c_int] * 7
        f("sisd_engine_new").restype = ctypes.c_void_p
        f("sisd_engine_new").argtypes = [ctypes.c_int] * 7
        f("engine_free").argtypes = [ctypes.c_void_p]
        f("graph_new").restype = ctypes.c_void_p
        f("graph_free").argtypes = [ctypes.c_void_p]
        f("graph_clear").argtypes = [ctypes.c_void_p]
        f("graph_num_nodes").restype = ctypes.c_uint32
        f("graph_num_nodes").argtypes = [ctypes.c_void_p]
        f("align").restype = ctypes.c_int
        f("align").argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                               ctypes.c_uint32, _c_i32p, ctypes.c_uint32, _c_i32p,
                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        f("graph_add_alignment").restype = ctypes.c_int
        f("graph_add_alignment").argtypes = [ctypes.c_void_p, _c_i32p, ctypes.c_uint32,
                                             ctypes.c_char_p, ctypes.c_uint32, _c_u32p,
                                             ctypes.c_uint32]
        f("graph_consensus").restype = ctypes.c_int
        f("graph_consensus").argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32]
        f("graph_msa").restype = ctypes.c_int
        f("graph_msa").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                   ctypes.c_uint64, _c_u32p, _c_u32p]
        f("graph_ranks").restype = ctypes.c_int
        f("graph_ranks").argtypes = [ctypes.c_void_p, _c_u32p, ctypes.c_uint32]

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def error(self):
        return self._fn("last_error")().decode()

    def engine(self, typ, m, n, g, e=None, q=None, c=None, sisd=False):
        """sisd=True: the reference's scalar engine (SisdAlignmentEngine)."""
        if e is None:
            e = g
        if q is None:
            q, c = g, e
        p = self._fn("sisd_engine_new" if sisd else "engine_new")(typ, m, n, g, e, q, c)
        if not p:
            raise ValueError(self.error())
        return p

    def free_engine(self, p):
        self._fn("engine_free")(p)

    def graph(self):
        return self._fn("graph_new")()

    def free_graph(self, p):
        self._fn("graph_free")(p)

    def num_nodes(self, gr):
        return self._fn("graph_num_nodes")(gr)

    def align(self, eng, gr, seq):
        """Returns (pairs as a list of (node, col), score or None, width bits)."""
        b = seq.encode() if isinstance(seq, str) else seq
        cap = self.num_nodes(gr) + len(b) + 8
        buf = (ctypes.c_int32 * (2 * cap))()
        score = ctypes.c_int32(0)
        has = ctypes.c_int(0)
        width = ctypes.c_int(0)
        n = self._fn("align")(eng, gr, b, len(b), buf, cap, ctypes.byref(score),
                              ctypes.byref(has), ctypes.byref(width))
        if n < 0:
            raise ValueError(self.error())
        pairs = [(buf[2 * k], buf[2 * k + 1]) for k in range(n)]
        return pairs, (score.value if has.value else None), width.value

    def add_alignment(self, gr, pairs, seq, weights=None):
        b = seq.encode() if isinstance(seq, str) else seq
        flat = (ctypes.c_int32 * max(1, 2 * len(pairs)))()
        for k, (a, c) in enumerate(pairs):
            flat[2 * k] = a
            flat[2 * k + 1] = c
        if weights is None:
            w, nw = None, 0
        else:
            w = (ctypes.c_uint32 * max(1, len(weights)))(*weights)
            nw = len(weights)
        rc = self._fn("graph_add_alignment")(gr, flat, len(pairs), b, len(b), w, nw)
        if rc != 0:
            raise ValueError(self.error())

    def consensus(self, gr):
        cap = 4 * self.num_nodes(gr) + 16
        buf = ctypes.create_string_buffer(cap)
        n = self._fn("graph_consensus")(gr, buf, cap)
        return buf.raw[:n].decode()

    def msa(self, gr, include_consensus=False):
        rows = ctypes.c_uint32(0)
        cols = ctypes.c_uint32(0)
        cap = 1 << 16
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = self._fn("graph_msa")(gr, int(include_consensus), buf, cap,
                                       ctypes.byref(rows), ctypes.byref(cols))
            if rc == -2:
                cap = rows.value * cols.value + 1
                continue
            break
        r, c = rows.value, cols.value
        raw = buf.raw
        return [raw[k * c:(k + 1) * c].decode() for k in range(r)]

    def ranks(self, gr):
        v = self.num_nodes(gr)
        buf = (ctypes.c_uint32 * max(1, v))()
        n = self._fn("graph_ranks")(gr, buf, v)
        return list(buf[:n])


def oracle():
    return Checker(ORACLE_SO, "oracle_")


def reference():
    """The compiled reference, or None where it was not built."""
    if not os.path.exists(REF_SO):
        return None
    return Checker(REF_SO, "ref_")


def run_window(chk, params, reads, weights=None, strand=False, calls=None, sisd=False):
    """Align reads in order and add each to one graph, as the reference's
    tests do (test/spoa_test.cpp:39-54).  With strand=True each read is also
    aligned reverse-complemented and the better strand is kept, as the
    reference's command line tool does (src/main.cpp:285-302).

    With sisd=True the engine is the reference's scalar one.

    Returns (consensus, msa rows, per-call records)."""
    eng = chk.engine(*params, sisd=sisd)
    gr = chk.graph()
    recs = [] if calls is None else calls
    try:
        for k, r in enumerate(reads):
            pairs, score, width = chk.align(eng, gr, r)
            rec = {"pairs": pairs, "score": score, "width": width, "read": k}
            seq = r
            if strand:
                rc = revcomp(r)
                pairs_r, score_r, width_r = chk.align(eng, gr, rc)
                rec["score_rc"] = score_r
                rec["pairs_rc"] = pairs_r
                s_f = score if score is not None else 0
                s_r = score_r if score_r is not None else 0
                if not (s_f >= s_r):
                    pairs, seq = pairs_r, rc
                rec["kept_forward"] = s_f >= s_r
            recs.append(rec)
            w = None if weights is None else weights[k]
            if strand and w is not None and seq is not r:
                w = list(reversed(w))
            chk.add_alignment(gr, pairs, seq, w)
        cons = chk.consensus(gr)
        msa = chk.msa(gr)
        return cons, msa, recs
    finally:
        chk.free_graph(gr)
        chk.free_engine(eng)


# The strand-ambiguous caller's reverse complement (src/main.cpp:286, 296
# call biosoup's Sequence::ReverseAndComplement, which the reference does
# not vendor): upper-case, then the Watson-Crick / IUPAC complement; every
# other character is kept.  The same rule as the product, the oracle and
# the reference shim; parity with biosoup is unpinned.
_SRC = b"ACGTURYKMBVDHSW"
_DST = b"TGCAAYRMKVBHDSW"
_COMP = bytes.maketrans(_SRC + _SRC.lower(), _DST + _DST)


def revcomp(s):
    b = s.encode() if isinstance(s, str) else s
    out = b.translate(_COMP)[::-1]
    return out.decode() if isinstance(s, str) else out
