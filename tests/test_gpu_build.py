"""Parity of the HIP graph build with the CPU oracle (byte for byte).

These are the parity tests proper: every case runs the GPU through the C ABI
(libgenomix_gx.so) and compares the record stream with the oracle's P = 1
restatement of the Hyracks build (edge lists in JDK-7 HashSet order), or,
in canonical mode, with the oracle after sorting each edge list.
"""

import glob
import os
import re
import struct

import numpy as np
import pytest

from oracle import oracle

from .conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "hyracks", "**", "*.expected.txt"), recursive=True))


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def gpu_records(gx, text, k, edge_order="jdk7", method=None):
    return gx.build_graph(text, k, edge_order=edge_order, method=method)[0]


def ora_records(text, k):
    return oracle.build_graph(text, k)[0]


def _show_diff(a, b):
    ta = oracle.dump_text(a).splitlines()
    tb = oracle.dump_text(b).splitlines()
    for i, (x, y) in enumerate(zip(ta, tb)):
        if x != y:
            return f"first differing record {i}:\n gpu: {x}\n ora: {y}"
    return f"record counts {len(ta)} vs {len(tb)}"


def assert_same(got, want):
    if got != want:
        pytest.fail(_show_diff(got, want))


@pytest.mark.parametrize("expected", FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def test_fixtures_byte_exact(gx, expected):
    text = open(expected.replace(".expected.txt", ".test.txt"), "rb").read()
    k = int(re.search(r"_k(\d+)\.", os.path.basename(expected)).group(1))
    assert_same(gpu_records(gx, text, k), ora_records(text, k))


def _sim(seed, G, L, cov, err, paired, n_mut=0):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=L, coverage=cov, error_rate=err, paired=paired,
                           insert_mean=3 * L, insert_sd=L // 4, seed=seed)
    return sim.simulate(spec)


@pytest.mark.parametrize("k", [3, 4, 5, 8, 11, 21, 31, 32, 33, 47, 55, 63])
@pytest.mark.parametrize("paired", [False, True], ids=["se", "pe"])
@pytest.mark.parametrize("method", ["sort", "hash"])
def test_random_reads_byte_exact(gx, k, paired, method):
    text = _sim(1000 + k, G=3000, L=max(k + 20, 70), cov=12, err=0.01, paired=paired)
    assert_same(gpu_records(gx, text, k, method=method), ora_records(text, k))


@pytest.mark.parametrize("k,G,cov", [(21, 20000, 100), (33, 8000, 150)])
def test_repeat_rich_long_replays_byte_exact(gx, k, G, cov):
    # repeat families at low divergence: k-mers with thousands of records and
    # branching neighbours (the long-run replay path)
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=100, coverage=cov, error_rate=0.01, paired=True,
                           insert_mean=300, insert_sd=30, repeat_frac=0.3, repeat_len=(200, 400),
                           repeat_div=0.01, seed=31337 + k)
    text = sim.simulate(spec)
    want = ora_records(text, k)
    for method in ("sort", "hash"):
        assert_same(gpu_records(gx, text, k, method=method), want)


@pytest.mark.parametrize("k", [4, 9, 31, 55])
def test_oversized_buckets_byte_exact(gx, k):
    # thousands of homopolymer and dinucleotide reads put far more than one
    # LDS bucket's worth of records on a few keys (global-sort fallback)
    rng = np.random.default_rng(k)
    lines = []
    for i in range(2500):
        unit = ["A", "C", "AC", "GT"][i % 4]
        s = (unit * 60)[:rng.integers(k + 5, 110)]
        s = "".join(c if rng.random() > 0.01 else rng.choice(list("ACGT")) for c in s)
        lines.append(f"{4 * i + 2}\t{s}\n")
    text = "".join(lines).encode()
    want = ora_records(text, k)
    for method in ("sort", "hash"):
        assert_same(gpu_records(gx, text, k, method=method), want)


def test_low_complexity_byte_exact(gx):
    # homopolymers and tandem repeats: self loops, palindromes (even k), many
    # records per node, lists of 4
    rng = np.random.default_rng(7)
    lines = []
    for i in range(300):
        unit = "".join(rng.choice(list("ACGT"), size=rng.integers(1, 5)))
        s = (unit * 80)[: rng.integers(20, 90)]
        s = "".join(c if rng.random() > 0.05 else rng.choice(list("ACGT")) for c in s)
        lines.append(f"{4 * i + 2}\t{s}\n")
    text = "".join(lines).encode()
    for k in (3, 4, 6, 9, 16):
        assert_same(gpu_records(gx, text, k), ora_records(text, k))


def test_line_format_edge_cases(gx):
    text = (b"1\tACGTACGTAC\r\n"            # CRLF
            b"2\tacgtNNACGT\tggccAATTGG\n"  # lowercase; invalid mate 0 still stored as mate
            b"3\t\tCCCCGGGGAT\n"            # empty mate 0
            b"4\tTTTTGGGGCC\t\t\n"          # trailing empty fields dropped
            b"5\tGATTACAGAT\r"              # lone CR terminator
            b"5\tGATTACAGAT\n"              # duplicate readId: TreeSet keeps one
            b"6\tACGTACGTAC")               # no final newline
    for k in (3, 5, 9):
        assert_same(gpu_records(gx, text, k), ora_records(text, k))


@pytest.mark.parametrize("bad", [b"1\n", b"1\tACGT\tACGT\tACGT\n", b"x\tACGTACGT\n", b"1\tACG\n", b"\n",
                                 b"2\tACGTAC\n536870912\tACGTACGT\n", b"1\tACGTAC\n\n2\tACGT\n"])
def test_errors_match_reference(gx, bad):
    with pytest.raises(oracle.OracleError):
        oracle.build_graph(bad, 3)
    with pytest.raises(gx.GxError) as e:
        gx.build_graph(bad, 3)
    assert e.value.code == gx.GX_EINPUT


def test_invalid_lines_without_valid_mates_pass(gx):
    text = b"-5\tNNNN\n7\tACGTTT\n"
    assert_same(gpu_records(gx, text, 3), ora_records(text, 3))


def _sorted_lists(records):
    out = []
    for line in oracle.dump_text(records).splitlines():
        f = line.split("\t")
        for i in (1, 2, 3, 4):
            name, v = f[i].split(":", 1)
            if v != "null":
                v = "[" + ",".join(sorted(v[1:-1].split(","))) + "]"
            f[i] = name + ":" + v
        out.append("\t".join(f))
    return out


@pytest.mark.parametrize("k", [3, 21, 55])
def test_canonical_edge_order_mode(gx, k):
    text = _sim(77 + k, G=2000, L=k + 40, cov=15, err=0.02, paired=False)
    got = gpu_records(gx, text, k, edge_order="canonical")
    assert _sorted_lists(got) == _sorted_lists(ora_records(text, k))
    # canonical mode: each list ascending by neighbour k-mer
    for line in oracle.dump_text(got).splitlines():
        for fld in line.split("\t")[1:5]:
            v = fld.split(":", 1)[1]
            if v != "null":
                items = v[1:-1].split(",")
                assert items == sorted(items)


def test_plain_cpu_workload_byte_exact(gx):
    """The 50 kbp / 100 bp / 30x / 0.5% error / k=21 workload, full size."""
    from genomix_amd import sim
    text = sim.simulate(sim.PLAIN_CPU_50KBP)
    assert_same(gpu_records(gx, text, 21), ora_records(text, 21))


def _iter_records(rec):
    off = 0
    while off < len(rec):
        n, kl = struct.unpack(">ii", rec[off:off + 8])
        yield rec[off + 8:off + 8 + kl], rec[off + 8 + kl:off + 8 + n]
        off += 8 + n


def test_plain_gpu_workload_invariants(gx):
    """5 Mbp, 150 bp PE, 50x, 1% error, k=31: size-independent properties at
    full size (the oracle is too slow here): sum of coverage = occurrences,
    one read head per valid mate, keys strictly increasing in KmerPointable
    order, deterministic bytes across runs."""
    from genomix_amd import sim
    text = sim.simulate(sim.PLAIN_GPU_5MBP)
    with gx.GraphBuilder(31) as b:
        b.add_reads(text)
        b.finish()
        rec = b.part_bytes(0)
        st = b.stats()
        b.reset()
        b.add_reads(text)
        b.finish()
        assert b.part_bytes(0) == rec
    nlines = text.count(b"\n")
    assert st["heads"] == 2 * nlines
    assert st["occurrences"] == 2 * nlines * (150 - 31 + 1)
    total_cov = 0
    prev = None
    nrec = 0
    for key, val in _iter_records(rec):
        kb = key[4:]
        rk = kb[::-1]
        assert prev is None or rk > prev
        prev = rk
        total_cov += struct.unpack(">f", val[-4:])[0]
        nrec += 1
    assert nrec == st["distinct"]
    assert total_cov == st["occurrences"]


def _digest_cases():
    import json
    path = os.path.join(GOLDEN, "digests", "plain_gpu_params.json")
    return json.load(open(path))["cases"]


@pytest.mark.parametrize("case", _digest_cases(), ids=lambda c: f"{c['genome_len'] // 1000}kbp")
def test_plain_gpu_params_byte_parity(gx, case):
    """The 5 Mbp workload's parameters (5% in 2 kb repeat copies at 1%, 150 bp
    PE, insert 500/50, 50x, 1% errors, k = 31) at 0.5 and 1 Mbp: graph and
    unitig records byte-identical to the oracle's, whose SHA-256 digests
    tests/golden/make_c2_digests.py recorded (the oracle needs minutes here)."""
    import dataclasses
    import hashlib
    from genomix_amd import sim
    spec = dataclasses.replace(sim.PLAIN_GPU_5MBP, genome_len=case["genome_len"], seed=case["seed"])
    text = sim.simulate(spec)
    assert hashlib.sha256(text).hexdigest() == case["text_sha256"], "simulator output changed"
    graph, merged = gx.build_and_merge(text, case["k"])
    assert len(graph) == case["graph_bytes"]
    assert hashlib.sha256(graph).hexdigest() == case["graph_sha256"]
    assert len(merged) == case["unitigs_bytes"]
    assert hashlib.sha256(merged).hexdigest() == case["unitigs_sha256"]
