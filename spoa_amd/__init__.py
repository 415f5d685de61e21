"""spoa_amd -- graph-to-sequence alignment (partial order alignment) on the
MI355X, behind spoa's own interface.

The C++ drop-in (``include/spoa/spoa.hpp``: ``spoa::AlignmentEngine``,
``spoa::Graph``) is compiled into ``libspoa_amd.so`` together with the HIP
kernels for gfx950; this module binds its C entry points
(``include/spoa_capi.h``) with ctypes and mirrors the reference interface:

* ``AlignmentEngine.create(type, m, n, g[, e[, q, c]])`` -- reference
  include/spoa/alignment_engine.hpp:35-55, same validation and messages;
* ``engine.align(sequence, graph)`` -- alignment_engine.hpp:61-69, returns
  the (node id, position) pairs and the score;
* ``Graph.add_alignment`` / ``generate_consensus`` /
  ``generate_multiple_sequence_alignment`` -- include/spoa/graph.hpp:134-171.

Beyond the reference: ``run_windows`` aligns many independent windows in
the batched form (one persistent launch per round) or in the plain per-call
form, and ``synth_windows`` produces the seeded synthetic windows the tests
and the benchmark use.  There is no CPU fallback: without the built library
or without a GPU every call raises.
"""
import ctypes
import os

import numpy as np

SW, NW, OV = 0, 1, 2
LINEAR, AFFINE, CONVEX = 0, 1, 2

HERE = os.path.dirname(os.path.abspath(__file__))
# The library built here; SPOA_AMD_LIBRARY, read when the library is first
# loaded, names another build of it (the row-phase timing build of
# `make -C spoa_amd phases`).
LIB_PATH = os.path.join(HERE, "libspoa_amd.so")

_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class HipStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("h2d_ms", ctypes.c_double),
                ("d2h_ms", ctypes.c_double), ("host_ms", ctypes.c_double),
                ("alloc_ms", ctypes.c_double), ("wait_ms", ctypes.c_double),
                ("unpack_ms", ctypes.c_double),
                ("launches", ctypes.c_uint32), ("jobs", ctypes.c_uint32),
                ("cells", ctypes.c_uint64), ("dp_bytes", ctypes.c_uint64),
                ("kernel_ms16", ctypes.c_double), ("kernel_ms32", ctypes.c_double),
                ("dp_bytes16", ctypes.c_uint64), ("dp_bytes32", ctypes.c_uint64),
                ("launches16", ctypes.c_uint32), ("launches32", ctypes.c_uint32),
                ("merge_ms", ctypes.c_double), ("merges", ctypes.c_uint32),
                ("graph_builds", ctypes.c_uint32), ("graph_updates", ctypes.c_uint32)]


class BatchOpts(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in
                ("type", "m", "n", "g", "e", "q", "c", "strand", "msa", "threads", "device", "overlap",
                 "graph", "resident", "sisd")]


class BatchStats(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("align_ms", ctypes.c_double),
                ("kernel_ms", ctypes.c_double), ("h2d_ms", ctypes.c_double),
                ("d2h_ms", ctypes.c_double), ("graph_ms", ctypes.c_double),
                ("finish_ms", ctypes.c_double), ("pack_ms", ctypes.c_double),
                ("alloc_ms", ctypes.c_double), ("wait_ms", ctypes.c_double),
                ("unpack_ms", ctypes.c_double), ("cells", ctypes.c_uint64),
                ("dp_bytes", ctypes.c_uint64), ("align_calls", ctypes.c_uint64),
                ("launches", ctypes.c_uint32), ("rounds", ctypes.c_uint32),
                ("kernel_ms16", ctypes.c_double), ("kernel_ms32", ctypes.c_double),
                ("dp_bytes16", ctypes.c_uint64), ("dp_bytes32", ctypes.c_uint64),
                ("launches16", ctypes.c_uint32), ("launches32", ctypes.c_uint32),
                ("graph_builds", ctypes.c_uint32), ("graph_updates", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeOpts(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_int32), ("device", ctypes.c_int32), ("max_active", ctypes.c_uint32),
                ("max_param_sets", ctypes.c_uint32), ("memory_fraction", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("submitted", "completed", "failed", "rounds", "launches",
                                                 "align_calls", "deferred", "cells", "dp_bytes")] + \
               [("kernel_ms", ctypes.c_double), ("max_active_seen", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SynthSpec(ctypes.Structure):
    _fields_ = [("template_min", ctypes.c_uint32), ("template_max", ctypes.c_uint32),
                ("reads_min", ctypes.c_uint32), ("reads_max", ctypes.c_uint32),
                ("read_len_min", ctypes.c_uint32), ("read_len_max", ctypes.c_uint32),
                ("err_min", ctypes.c_double), ("err_max", ctypes.c_double),
                ("revcomp_odd", ctypes.c_uint32), ("log_uniform_len", ctypes.c_uint32)]


def lib():
    """The native library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.environ.get("SPOA_AMD_LIBRARY") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            "spoa_amd: %s is missing; build it with __graft_entry__.build() "
            "(make -C spoa_amd)" % path)
    L = ctypes.CDLL(path)
    L.spoa_last_error.restype = ctypes.c_char_p
    L.spoa_engine_new.restype = ctypes.c_void_p
    L.spoa_engine_new.argtypes = [ctypes.c_int] * 7
    L.spoa_sisd_engine_new.restype = ctypes.c_void_p
    L.spoa_sisd_engine_new.argtypes = [ctypes.c_int] * 7
    L.spoa_hip_row_phases.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32,
                                      ctypes.c_int32]
    L.spoa_hip_last_error.restype = ctypes.c_char_p
    L.spoa_engine_free.argtypes = [ctypes.c_void_p]
    L.spoa_engine_subtype.argtypes = [ctypes.c_void_p]
    L.spoa_engine_prealloc.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    L.spoa_engine_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(HipStats)]
    L.spoa_engine_device.restype = ctypes.c_void_p
    L.spoa_engine_device.argtypes = [ctypes.c_void_p]
    L.spoa_align.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32,
                             _i32p, ctypes.c_uint32, _i32p, ctypes.POINTER(ctypes.c_int),
                             ctypes.POINTER(ctypes.c_int)]
    L.spoa_graph_new.restype = ctypes.c_void_p
    L.spoa_graph_free.argtypes = [ctypes.c_void_p]
    L.spoa_graph_clear.argtypes = [ctypes.c_void_p]
    L.spoa_graph_num_nodes.restype = ctypes.c_uint32
    L.spoa_graph_num_nodes.argtypes = [ctypes.c_void_p]
    L.spoa_graph_add_alignment.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_uint32, ctypes.c_char_p,
                                           ctypes.c_uint32, _u32p, ctypes.c_uint32]
    L.spoa_graph_consensus.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32]
    L.spoa_graph_msa.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64,
                                 _u32p, _u32p]
    L.spoa_graph_ranks.argtypes = [ctypes.c_void_p, _u32p, ctypes.c_uint32]
    L.spoa_batch_run.restype = ctypes.c_void_p
    L.spoa_batch_run.argtypes = [ctypes.c_int, ctypes.POINTER(BatchOpts), ctypes.c_uint32,
                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p]
    L.spoa_batch_free.argtypes = [ctypes.c_void_p]
    L.spoa_batch_context_new.restype = ctypes.c_void_p
    L.spoa_batch_context_new.argtypes = [ctypes.POINTER(BatchOpts)]
    L.spoa_batch_context_free.argtypes = [ctypes.c_void_p]
    L.spoa_batch_context_device.restype = ctypes.c_void_p
    L.spoa_batch_context_device.argtypes = [ctypes.c_void_p]
    L.spoa_batch_context_half_device.restype = ctypes.c_void_p
    L.spoa_batch_context_half_device.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    L.spoa_batch_consensus_all.restype = ctypes.c_int64
    L.spoa_batch_consensus_all.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_void_p]
    L.spoa_batch_context_set_config.restype = ctypes.c_int
    L.spoa_batch_context_set_config.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_int32,
                                                ctypes.c_int32]
    L.spoa_batch_context_set_config_split.restype = ctypes.c_int
    L.spoa_batch_context_set_config_split.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32,
                                                      ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.spoa_batch_context_autotune.restype = ctypes.c_int
    L.spoa_batch_context_autotune.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_uint32]
    L.spoa_batch_context_run.restype = ctypes.c_void_p
    L.spoa_batch_context_run.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_char_p, ctypes.c_char_p]
    L.spoa_batch_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchStats)]
    L.spoa_batch_consensus.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_uint32]
    L.spoa_batch_msa.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_uint64,
                                 _u32p, _u32p]
    L.spoa_batch_reversed.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_uint32]
    L.spoa_synth_new.restype = ctypes.c_void_p
    L.spoa_synth_new.argtypes = [ctypes.POINTER(SynthSpec), ctypes.c_uint64, ctypes.c_uint32]
    L.spoa_synth_free.argtypes = [ctypes.c_void_p]
    L.spoa_synth_sizes.argtypes = [ctypes.c_void_p, _u32p, _u64p, _u64p, _u64p]
    L.spoa_synth_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p,
                                  ctypes.c_void_p, ctypes.c_char_p]
    L.spoa_server_new.restype = ctypes.c_void_p
    L.spoa_server_new.argtypes = [ctypes.POINTER(ServeOpts)]
    L.spoa_server_free.argtypes = [ctypes.c_void_p]
    L.spoa_server_submit.restype = ctypes.c_uint64
    L.spoa_server_submit.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 7 + [
        ctypes.c_uint32, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.spoa_server_wait.restype = ctypes.c_void_p
    L.spoa_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_double]
    L.spoa_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServeStats)]
    L.spoa_server_now_ms.restype = ctypes.c_double
    L.spoa_server_now_ms.argtypes = [ctypes.c_void_p]
    L.spoa_serve_result_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)] + \
        [ctypes.POINTER(ctypes.c_double)] * 3 + [_u32p]
    L.spoa_serve_result_lane.argtypes = [ctypes.c_void_p]
    L.spoa_serve_result_error.restype = ctypes.c_char_p
    L.spoa_serve_result_error.argtypes = [ctypes.c_void_p]
    L.spoa_serve_result_consensus.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32]
    L.spoa_serve_result_msa.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, _u32p, _u32p]
    L.spoa_serve_result_free.argtypes = [ctypes.c_void_p]
    # C ABI of the kernels (include/spoa_hip.h), for tuning and budgets
    L.spoa_hip_last_error.restype = ctypes.c_char_p
    L.spoa_hip_set_config.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                      ctypes.c_int32, ctypes.c_int32]
    L.spoa_hip_config.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                                  _i32p, _i32p]
    L.spoa_hip_set_config_split.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.spoa_hip_config_split.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int32,
                                        _i32p, _i32p, _i32p]
    L.spoa_hip_set_wait_limit.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.spoa_hip_set_memory_budget.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
    L.spoa_hip_memory_budget.restype = ctypes.c_uint64
    L.spoa_hip_memory_budget.argtypes = [ctypes.c_void_p]
    L.spoa_hip_memory_in_use.restype = ctypes.c_uint64
    L.spoa_hip_memory_in_use.argtypes = [ctypes.c_void_p]
    L.spoa_reverse_complement.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_char_p]
    L.spoa_hip_job_footprint.restype = ctypes.c_uint64
    L.spoa_hip_job_footprint.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32]
    _lib = L
    return L


def _error():
    return lib().spoa_last_error().decode()


def _b(s):
    return s.encode() if isinstance(s, str) else bytes(s)


def reverse_complement(sequence):
    """The reverse complement that strand-ambiguous windows align (upper-case,
    then the Watson-Crick / IUPAC complement; other bytes kept)."""
    s = _b(sequence)
    out = ctypes.create_string_buffer(max(1, len(s)))
    lib().spoa_reverse_complement(s, len(s), out)
    r = out.raw[:len(s)]
    return r.decode() if isinstance(sequence, str) else r


class Graph:
    """spoa::Graph (reference include/spoa/graph.hpp:26-318)."""

    def __init__(self):
        self._h = lib().spoa_graph_new()

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.spoa_graph_free(self._h)
            self._h = None

    @property
    def num_nodes(self):
        return lib().spoa_graph_num_nodes(self._h)

    def clear(self):
        lib().spoa_graph_clear(self._h)

    def add_alignment(self, alignment, sequence, weights=1):
        """Graph::AddAlignment: weights is an int (every base), a list of
        per-base weights, or a Phred+33 quality string (graph.hpp:134-162)."""
        s = _b(sequence)
        flat = (ctypes.c_int32 * max(1, 2 * len(alignment)))()
        for k, (a, c) in enumerate(alignment):
            flat[2 * k] = a
            flat[2 * k + 1] = c
        if isinstance(weights, (str, bytes)):
            w = [x - 33 for x in _b(weights)]
        elif isinstance(weights, int):
            w = [weights] * len(s)
        else:
            w = list(weights)
        wa = (ctypes.c_uint32 * max(1, len(w)))(*w)
        if lib().spoa_graph_add_alignment(self._h, flat, len(alignment), s, len(s), wa, len(w)) != 0:
            raise ValueError(_error())

    def generate_consensus(self):
        cap = 4 * self.num_nodes + 16
        buf = ctypes.create_string_buffer(cap)
        n = lib().spoa_graph_consensus(self._h, buf, cap)
        return buf.raw[:n].decode()

    def generate_multiple_sequence_alignment(self, include_consensus=False):
        rows, cols = ctypes.c_uint32(0), ctypes.c_uint32(0)
        cap = 1 << 16
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = lib().spoa_graph_msa(self._h, int(include_consensus), buf, cap,
                                      ctypes.byref(rows), ctypes.byref(cols))
            if rc == -2:
                cap = rows.value * cols.value + 1
                continue
            break
        raw = buf.raw
        c = cols.value
        return [raw[k * c:(k + 1) * c].decode() for k in range(rows.value)]

    def rank_to_node(self):
        v = self.num_nodes
        buf = (ctypes.c_uint32 * max(1, v))()
        n = lib().spoa_graph_ranks(self._h, buf, v)
        return list(buf[:n])


class AlignmentEngine:
    """spoa::AlignmentEngine (reference include/spoa/alignment_engine.hpp:32-96)
    whose Align runs on the GPU."""

    def __init__(self, handle):
        self._h = handle
        self.last_width = 0

    @staticmethod
    def create(type, m, n, g, e=None, q=None, c=None, sisd=False):
        """sisd=True: the reference's scalar engine's results
        (SisdAlignmentEngine, what a reference built without SIMD runs)."""
        if e is None:
            e = g
        if q is None:
            q, c = g, e
        h = (lib().spoa_sisd_engine_new if sisd else lib().spoa_engine_new)(type, m, n, g, e, q, c)
        if not h:
            raise ValueError(_error())
        return AlignmentEngine(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.spoa_engine_free(self._h)
            self._h = None

    @property
    def subtype(self):
        return lib().spoa_engine_subtype(self._h)

    @property
    def device(self):
        return lib().spoa_engine_device(self._h)

    def prealloc(self, max_sequence_len, alphabet_size):
        if lib().spoa_engine_prealloc(self._h, max_sequence_len, alphabet_size) != 0:
            raise ValueError(_error())

    def align(self, sequence, graph):
        """Returns (alignment, score); score is None where the reference
        leaves *score unwritten (empty graph or read, no end cell)."""
        s = _b(sequence)
        cap = graph.num_nodes + len(s) + 8
        buf = (ctypes.c_int32 * (2 * cap))()
        score = ctypes.c_int32(0)
        has = ctypes.c_int(0)
        width = ctypes.c_int(0)
        n = lib().spoa_align(self._h, graph._h, s, len(s), buf, cap, ctypes.byref(score),
                             ctypes.byref(has), ctypes.byref(width))
        if n < 0:
            raise ValueError(_error())
        self.last_width = width.value
        return [(buf[2 * k], buf[2 * k + 1]) for k in range(n)], (score.value if has.value else None)

    def stats(self):
        st = HipStats()
        lib().spoa_engine_stats(self._h, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_}


# ------------------------------------------------------------- windows --

class Windows:
    """A flattened set of windows: reads of window w are
    ``seqs[read_off[k]:read_off[k+1]]`` for k in
    ``range(win_read_off[w], win_read_off[w+1])``."""

    def __init__(self, win_read_off, read_off, seqs, quals=None, templates=None):
        self.win_read_off = np.ascontiguousarray(win_read_off, dtype=np.uint32)
        self.read_off = np.ascontiguousarray(read_off, dtype=np.uint64)
        self.seqs = bytes(seqs)
        self.quals = None if quals is None else bytes(quals)
        self.templates = templates

    @staticmethod
    def from_lists(windows, qualities=None):
        wo, ro, parts, qparts = [0], [0], [], []
        for w, reads in enumerate(windows):
            for k, r in enumerate(reads):
                b = _b(r)
                parts.append(b)
                ro.append(ro[-1] + len(b))
                if qualities is not None:
                    qparts.append(_b(qualities[w][k]))
            wo.append(len(ro) - 1)
        return Windows(wo, ro, b"".join(parts), b"".join(qparts) if qualities is not None else None)

    @property
    def n_windows(self):
        return len(self.win_read_off) - 1

    def max_read_len(self):
        """Length of the longest read of every window."""
        return int(np.diff(self.read_off).max()) if len(self.read_off) > 1 else 0

    def reads(self, w):
        lo, hi = int(self.win_read_off[w]), int(self.win_read_off[w + 1])
        return [self.seqs[int(self.read_off[k]):int(self.read_off[k + 1])].decode() for k in range(lo, hi)]

    def qualities(self, w):
        if self.quals is None:
            return None
        lo, hi = int(self.win_read_off[w]), int(self.win_read_off[w + 1])
        return [self.quals[int(self.read_off[k]):int(self.read_off[k + 1])].decode() for k in range(lo, hi)]

    def subset(self, idx):
        return Windows.from_lists([self.reads(w) for w in idx],
                                  None if self.quals is None else [self.qualities(w) for w in idx])


class _MsaView:
    """Every window's MSA, read from the native results on first access of
    each window (rows of equal width, as Graph::GenerateMultipleSequenceAlignment
    returns them)."""

    def __init__(self, owner, n, enabled):
        self._owner = owner
        self._n = n if enabled else 0
        self._cache = {}

    def __len__(self):
        return self._n

    def __getitem__(self, w):
        if isinstance(w, slice):
            return [self[i] for i in range(*w.indices(self._n))]
        if w < 0:
            w += self._n
        if not 0 <= w < self._n:
            raise IndexError(w)
        if w not in self._cache:
            self._cache[w] = self._owner._msa(w)
        return self._cache[w]

    def __iter__(self):
        return (self[i] for i in range(self._n))

    def __eq__(self, other):
        return len(self) == len(other) and all(a == b for a, b in zip(self, other))


class WindowResults:
    """Results of a window run: consensus strings (a list), the MSAs (read
    from the native results window by window, on access), the strand flags
    and the run's measurements."""

    def __init__(self, h, windows, msa, strand):
        self._h = h
        try:
            self._load(windows, msa, strand)
        except BaseException:
            self._h = None
            lib().spoa_batch_free(h)
            raise

    def _load(self, windows, msa, strand):
        h = self._h
        n = windows.n_windows
        st = BatchStats()
        lib().spoa_batch_get_stats(h, ctypes.byref(st))
        self.stats = st.as_dict()
        offs = np.zeros(n + 1, dtype=np.uint64)
        cap = 1 << 20
        while True:
            buf = ctypes.create_string_buffer(cap)
            k = lib().spoa_batch_consensus_all(h, buf, cap, offs.ctypes.data)
            if k == -2:
                cap = int(offs[n]) + 1
                continue
            break
        raw = buf.raw[:k].decode()
        o = offs.tolist()
        self.consensus = [raw[o[w]:o[w + 1]] for w in range(n)]
        self.msa = _MsaView(self, n, msa)
        self.reversed = []
        if strand:
            for w in range(n):
                nr = int(windows.win_read_off[w + 1] - windows.win_read_off[w])
                rb = ctypes.create_string_buffer(max(1, nr))
                k = lib().spoa_batch_reversed(h, w, rb, nr)
                self.reversed.append(list(rb.raw[:k]))

    def _msa(self, w):
        rows, cols = ctypes.c_uint32(0), ctypes.c_uint32(0)
        cap = 1 << 16
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = lib().spoa_batch_msa(self._h, w, buf, cap, ctypes.byref(rows), ctypes.byref(cols))
            if rc == -2:
                cap = rows.value * cols.value + 1
                continue
            break
        cc = cols.value
        raw = buf.raw
        return [raw[i * cc:(i + 1) * cc].decode() for i in range(rows.value)]

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.spoa_batch_free(self._h)
            self._h = None


def _opts(type, m, n, g, e, q, c, strand, msa, threads, device, overlap=False, graph=False, resident=False,
          sisd=False):
    if e is None:
        e = g
    if q is None:
        q, c = g, e
    return BatchOpts(type=type, m=m, n=n, g=g, e=e, q=q, c=c, strand=int(strand), msa=int(msa),
                     threads=threads, device=device, overlap=int(overlap), graph=int(graph),
                     resident=int(resident), sisd=int(sisd))


class BatchContext:
    """A reusable batched-form engine (device buffers and host threads kept
    between runs) for one set of scoring options."""

    def __init__(self, type, m, n, g, e=None, q=None, c=None, strand=False, msa=True, threads=0, device=0,
                 overlap=False, graph=False, resident=False, sisd=False):
        self.opts = _opts(type, m, n, g, e, q, c, strand, msa, threads, device, overlap, graph, resident, sisd)
        self._h = lib().spoa_batch_context_new(ctypes.byref(self.opts))
        if not self._h:
            raise ValueError(_error())

    def __del__(self):
        self.close()

    def close(self):
        """Releases the context's engines, device memory and host threads
        now (the context is unusable afterwards)."""
        if getattr(self, "_h", None) and _lib is not None:
            _lib.spoa_batch_context_free(self._h)
            self._h = None

    @property
    def device(self):
        return lib().spoa_batch_context_device(self._h)

    def device_of(self, half):
        """The C ABI engine of one half of the windows (0 or 1)."""
        return lib().spoa_batch_context_half_device(self._h, half)

    def autotune(self, windows, rounds=8, cache=None):
        """Picks the launch configuration (columns per lane, waves) of each
        score width by timing candidates on these windows (their first
        rounds, or the whole batch for a width those rounds do not reach);
        returns [{width, cols_per_lane, waves, cached, whole_batch, kernel_ms}]."""
        if not isinstance(windows, Windows):
            windows = Windows.from_lists(windows)
        out = np.zeros(8, dtype=np.int32)
        ms = np.zeros(2, dtype=np.float64)
        n = lib().spoa_batch_context_autotune(self._h, windows.n_windows, windows.win_read_off.ctypes.data,
                                              windows.read_off.ctypes.data, windows.seqs, rounds,
                                              _b(cache) if cache else None, out.ctypes.data, ms.ctypes.data, 2)
        if n < 0:
            raise ValueError(_error())
        return [{"width": int(out[4 * i]), "cols_per_lane": int(out[4 * i + 1]), "waves": int(out[4 * i + 2]),
                 "cached": bool(out[4 * i + 3] & 1), "whole_batch": bool(out[4 * i + 3] & 2),
                 "kernel_ms": float(ms[i])} for i in range(n)]

    def set_config(self, width_bits, cols_per_lane, waves, max_width=0xFFFFFFFF, split=False):
        """Launch configuration of the batched launches for rows up to
        max_width columns at one score width; cols_per_lane 0 clears.  With
        split, a row of several tiles runs each tile on its own workgroup."""
        if lib().spoa_batch_context_set_config_split(self._h, width_bits, max_width, cols_per_lane, waves,
                                                     1 if split else 0) != 0:
            raise ValueError(_error())

    def run(self, windows):
        if not isinstance(windows, Windows):
            windows = Windows.from_lists(windows)
        h = lib().spoa_batch_context_run(self._h, windows.n_windows, windows.win_read_off.ctypes.data,
                                         windows.read_off.ctypes.data, windows.seqs, windows.quals)
        return _collect(h, windows, bool(self.opts.msa), bool(self.opts.strand))


def run_windows(windows, type, m, n, g, e=None, q=None, c=None, mode="batched", strand=False,
                msa=True, threads=0, device=0, sisd=False):
    """Aligns every read of every window in order into the window's graph,
    then builds its consensus (and MSA).  mode 'batched': one launch per
    round over all windows, graphs merged on the host; 'resident': the same
    launches with every window's graph kept, merged, sorted and summarised
    on the device; 'per_call': the drop-in Align, one read at a time.
    sisd=True: the reference's scalar engine's alignments."""
    if not isinstance(windows, Windows):
        windows = Windows.from_lists(windows)
    o = _opts(type, m, n, g, e, q, c, strand, msa, threads, device, resident=mode == "resident", sisd=sisd)
    md = 1 if mode == "per_call" else 0
    h = lib().spoa_batch_run(md, ctypes.byref(o), windows.n_windows,
                             windows.win_read_off.ctypes.data, windows.read_off.ctypes.data,
                             windows.seqs, windows.quals)
    return _collect(h, windows, msa, strand)


def _collect(h, windows, msa, strand):
    if not h:
        raise ValueError(_error())
    return WindowResults(h, windows, msa, strand)


class ServeResult:
    def __init__(self, status, error, consensus, msa, submit_ms, admit_ms, done_ms, rounds, lane=0):
        self.lane = lane  # size-class lane by first-read length (< 700, < 2,000, < 5,000, longer)
        self.status = status
        self.error = error
        self.consensus = consensus
        self.msa = msa
        self.submit_ms = submit_ms
        self.admit_ms = admit_ms
        self.done_ms = done_ms
        self.rounds = rounds

    @property
    def latency_ms(self):
        return self.done_ms - self.submit_ms


class Server:
    """The serving loop on one GPU: consensus jobs (a window of reads and its
    scoring parameters) are admitted as slots and device memory free up and
    advance one read per round in continuous batches."""

    def __init__(self, threads=0, device=0, max_active=4096, max_param_sets=2, memory_fraction=0.6):
        o = ServeOpts(threads, device, max_active, max_param_sets, memory_fraction)
        self._h = lib().spoa_server_new(ctypes.byref(o))
        if not self._h:
            raise ValueError(_error())

    def close(self):
        """Finishes every submitted job, then stops the server."""
        if getattr(self, "_h", None) and _lib is not None:
            _lib.spoa_server_free(self._h)
            self._h = None

    __del__ = close

    def submit(self, reads, params, qualities=None, msa=True):
        t, m, n, g, e, q, c = tuple(params) + (None,) * (7 - len(params))
        e = g if e is None else e
        if q is None:
            q, c = g, e
        enc = [_b(r) for r in reads]
        off = np.zeros(len(enc) + 1, dtype=np.uint64)
        off[1:] = np.cumsum([len(x) for x in enc]) if enc else []
        quals = b"".join(_b(x) for x in qualities) if qualities is not None else None
        jid = lib().spoa_server_submit(self._h, t, m, n, g, e, q, c, len(enc), off.ctypes.data, b"".join(enc),
                                       quals, int(msa))
        if jid == 0:
            raise ValueError(_error())
        return jid

    def now_ms(self):
        return lib().spoa_server_now_ms(self._h)

    def wait(self, jid, timeout=-1.0):
        """The job's ServeResult, or None on timeout."""
        r = lib().spoa_server_wait(self._h, jid, timeout)
        if not r:
            return None
        try:
            st = ctypes.c_int(0)
            t = [ctypes.c_double(0) for _ in range(3)]
            rounds = ctypes.c_uint32(0)
            lib().spoa_serve_result_info(r, ctypes.byref(st), ctypes.byref(t[0]), ctypes.byref(t[1]),
                                         ctypes.byref(t[2]), ctypes.byref(rounds))
            err = lib().spoa_serve_result_error(r).decode()
            cap = 1 << 16
            while True:
                buf = ctypes.create_string_buffer(cap)
                k = lib().spoa_serve_result_consensus(r, buf, cap)
                if k == -2:
                    cap *= 4
                    continue
                break
            cons = buf.raw[:k].decode()
            rows, cols = ctypes.c_uint32(0), ctypes.c_uint32(0)
            cap = 1 << 16
            while True:
                mb = ctypes.create_string_buffer(cap)
                rc = lib().spoa_serve_result_msa(r, mb, cap, ctypes.byref(rows), ctypes.byref(cols))
                if rc == -2:
                    cap = rows.value * cols.value + 1
                    continue
                break
            raw = mb.raw
            msa = [raw[i * cols.value:(i + 1) * cols.value].decode() for i in range(rows.value)]
            return ServeResult(st.value, err, cons, msa, t[0].value, t[1].value, t[2].value, rounds.value,
                               lib().spoa_serve_result_lane(r))
        finally:
            lib().spoa_serve_result_free(r)

    def stats(self):
        st = ServeStats()
        lib().spoa_server_stats(self._h, ctypes.byref(st))
        return st.as_dict()


def serve_trace(jobs, arrivals_s, server=None, msa=True, **server_args):
    """Replays an arrival trace: job i (reads, params) is submitted at
    arrivals_s[i] seconds after the start; returns (results in job order,
    wall seconds from the first arrival to the last completion, stats)."""
    import threading
    import time
    own = server is None
    srv = Server(**server_args) if own else server
    ids = [None] * len(jobs)
    err = []

    def feeder():
        t0 = time.perf_counter()
        try:
            for i, (reads, params) in enumerate(jobs):
                dt = arrivals_s[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids[i] = srv.submit(reads, params, msa=msa)
        except BaseException as ex:  # surfaced below
            err.append(ex)

    t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    if err:
        raise err[0]
    results = [srv.wait(j) for j in ids]
    wall = time.perf_counter() - t0
    stats = srv.stats()
    if own:
        srv.close()
    return results, wall, stats


def synth_windows(seed, n_windows, template=(500, 500), reads=(20, 20), read_len=(0, 0),
                  err=(0.1, 0.1), revcomp_odd=False, log_uniform_len=False):
    """Seeded synthetic windows: uniform ACGT templates, reads as template
    substrings (length clipped to the template) or whole templates, per-base
    errors split evenly into substitutions, insertions and deletions."""
    sp = SynthSpec(template[0], template[1], reads[0], reads[1], read_len[0], read_len[1],
                   err[0], err[1], int(revcomp_odd), int(log_uniform_len))
    h = lib().spoa_synth_new(ctypes.byref(sp), seed, n_windows)
    if not h:
        raise ValueError(_error())
    try:
        nw, nr = ctypes.c_uint32(0), ctypes.c_uint64(0)
        sb, tb = ctypes.c_uint64(0), ctypes.c_uint64(0)
        lib().spoa_synth_sizes(h, ctypes.byref(nw), ctypes.byref(nr), ctypes.byref(sb), ctypes.byref(tb))
        wo = np.zeros(nw.value + 1, dtype=np.uint32)
        ro = np.zeros(nr.value + 1, dtype=np.uint64)
        to = np.zeros(nw.value + 1, dtype=np.uint64)
        seqs = ctypes.create_string_buffer(max(1, sb.value))
        tpl = ctypes.create_string_buffer(max(1, tb.value))
        lib().spoa_synth_copy(h, wo.ctypes.data, ro.ctypes.data, seqs, to.ctypes.data, tpl)
        templates = [tpl.raw[int(to[w]):int(to[w + 1])].decode() for w in range(nw.value)]
        return Windows(wo, ro, seqs.raw[:sb.value], None, templates)
    finally:
        lib().spoa_synth_free(h)


# The five workloads (window shapes and scoring) the benchmark and tests use.
CONFIGS = {
    1: dict(name="cfg1_nw_linear_500bp_20reads", n_windows=1, template=(500, 500), reads=(20, 20),
            read_len=(0, 0), err=(0.10, 0.10), params=(NW, 3, -5, -3)),
    2: dict(name="cfg2_sw_convex_1.5kb_30reads", n_windows=1, template=(1500, 1500), reads=(30, 30),
            read_len=(1000, 2000), err=(0.15, 0.15), params=(SW, 5, -4, -8, -6, -10, -4)),
    3: dict(name="cfg3_ov_affine_500bp_4096windows", n_windows=4096, template=(500, 500),
            reads=(10, 50), read_len=(400, 600), err=(0.12, 0.12), params=(OV, 5, -4, -8, -6)),
    4: dict(name="cfg4_nw_convex_5-10kb_256windows_strand", n_windows=256, template=(5000, 10000),
            reads=(8, 32), read_len=(0, 0), err=(0.10, 0.10), revcomp_odd=True, strand=True,
            params=(NW, 5, -4, -8, -6, -10, -4)),
    # serving: jobs of both the cfg-2 and the cfg-3 scoring, reads spanning a
    # template of log-uniform length; job i takes params_mix[i % 2]
    5: dict(name="cfg5_serving_mix_200bp-10kb", n_windows=512, template=(200, 10000), reads=(5, 100),
            read_len=(0, 0), err=(0.08, 0.15), log_uniform_len=True,
            params=(SW, 5, -4, -8, -6, -10, -4),
            params_mix=((SW, 5, -4, -8, -6, -10, -4), (OV, 5, -4, -8, -6))),
}


def config_windows(cfg, seed, n_windows=None):
    c = CONFIGS[cfg]
    return synth_windows(seed, c["n_windows"] if n_windows is None else n_windows,
                         template=c["template"], reads=c["reads"], read_len=c["read_len"],
                         err=c["err"], revcomp_odd=c.get("revcomp_odd", False),
                         log_uniform_len=c.get("log_uniform_len", False))


def serving_jobs(seed, n_jobs, rate_per_s, cfg=5, template=None, reads=None):
    """Seeded serving workload: n_jobs (reads, params) windows of the serving
    mix and their Poisson arrival times (seconds) at rate_per_s."""
    import random
    c = CONFIGS[cfg]
    ws = synth_windows(seed, n_jobs, template=template or c["template"], reads=reads or c["reads"],
                       read_len=c["read_len"], err=c["err"], log_uniform_len=c.get("log_uniform_len", False))
    rng = random.Random(seed)
    t, arrivals = 0.0, []
    for _ in range(n_jobs):
        t += rng.expovariate(rate_per_s)
        arrivals.append(t)
    jobs = [(ws.reads(i), c["params_mix"][i % len(c["params_mix"])]) for i in range(n_jobs)]
    return jobs, arrivals
