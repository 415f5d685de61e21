"""The command-line tool (spoa_amd/spoa, the reference's src/main.cpp on
the GPU engine) on the reference's test data (gzip FASTQ with qualities):
consensus, MSA, GFA and DOT outputs equal the compiled reference's, built
from the same reads in the same order (its GFA printer restated in the
reference shim), in all three alignment modes and in strand mode."""
import ctypes
import os
import subprocess

import pytest

import common
import oracle_lib as ol

pytestmark = pytest.mark.gpu

TOOL = os.path.join(common.ROOT, "spoa_amd", "spoa")
SAMPLE = os.path.join(common.GOLDEN, "sample.fastq.gz")


def _run(args):
    out = subprocess.run([TOOL] + args + [SAMPLE], capture_output=True, timeout=300)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    return out.stdout.decode()


def _expected(params, strand, tmp_path):
    """(consensus, MSA rows with consensus, GFA with consensus, DOT) of the
    reference for the sample reads."""
    ref = ol.reference()
    data = common.sample_reads()
    eng = ref.engine(*params)
    g = ref.graph()
    rev = []
    for s, q in data:
        w = [ord(c) - 33 for c in q]
        pairs, score, _ = ref.align(eng, g, s)
        seq = s
        if strand:
            rc = ol.revcomp(s)
            pr, sr, _ = ref.align(eng, g, rc)
            if (score or 0) >= (sr or 0):
                rev.append(0)
            else:
                pairs, seq, w = pr, rc, w[::-1]
                rev.append(1)
        ref.add_alignment(g, pairs, seq, w)
    cons = ref.consensus(g)
    msa = ref.msa(g, include_consensus=True)
    f = ref.lib.ref_graph_gfa
    f.restype = ctypes.c_int64
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_int,
                  ctypes.c_char_p, ctypes.c_uint64]
    headers = b"".join(str(i).encode() + b"\0" for i in range(len(data)))
    buf = ctypes.create_string_buffer(1 << 22)
    k = f(g, headers, bytes(rev) if strand else None, len(rev), 1, buf, 1 << 22)
    gfa = buf.raw[:k].decode()
    dot = tmp_path / "ref.dot"
    ref.lib.ref_graph_print_dot.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
    ref.lib.ref_graph_print_dot(g, str(dot).encode())
    ref.free_graph(g)
    ref.free_engine(eng)
    return cons, msa, gfa, dot.read_text()


@pytest.mark.skipif(ol.reference() is None, reason="needs the compiled reference")
@pytest.mark.parametrize("algo,gaps,strand", [(0, [], False), (1, ["-g", "-8", "-e", "-8"], False),
                                              (2, ["-e", "-6", "-q", "-8", "-c", "-6"], True)])
def test_cli_outputs_equal_the_reference(tmp_path, algo, gaps, strand):
    opts = {"-m": 5, "-n": -4, "-g": -8, "-e": -6, "-q": -10, "-c": -4}
    for k in range(0, len(gaps), 2):
        opts[gaps[k]] = int(gaps[k + 1])
    params = (algo, opts["-m"], opts["-n"], opts["-g"], opts["-e"], opts["-q"], opts["-c"])
    cons, msa, gfa, dot = _expected(params, strand, tmp_path)
    args = ["-l", str(algo)] + gaps + (["-s"] if strand else [])
    out0 = _run(args)
    assert out0 == ">Consensus LN:i:%d\n%s\n" % (len(cons), cons)
    names = [str(i) for i in range(len(msa) - 1)] + ["Consensus"]
    path = tmp_path / "tool.dot"
    out = _run(args + ["-r", "2", "-r", "4", "-d", str(path)])
    want = "".join(">%s\n%s\n" % (n, row) for n, row in zip(names, msa)) + gfa
    assert out == want
    assert path.read_text() == dot
