"""BUBBLE then MERGE on the device (include/genomix_clean.h
gx_clean_bubbles) against the oracle's restatement, byte for byte."""

import math

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def _sim(seed, G=40000, L=100, cov=20, paired=True, err=0.01):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=L, coverage=cov, error_rate=err, paired=paired,
                           insert_mean=3 * L, insert_sd=L // 4, repeat_frac=0.05, repeat_len=(300, 900),
                           repeat_div=0.01, seed=seed)
    return sim.simulate(spec)


def _readids(seqs):
    return b"".join(b"%d\t%s\n" % (4 * i + 2, s.encode()) for i, s in enumerate(seqs))


def _rc(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def _alleles(seed, G, L, alt_of, frac=0.3):
    rng = np.random.default_rng(seed)
    genome = "".join("ACGT"[x] for x in rng.integers(0, 4, G))
    alt = alt_of(genome)
    reads = []
    for i, s in enumerate(range(0, G - L + 1, 2)):
        r = (alt if (i % 10) < int(frac * 10) else genome)[s:s + L]
        reads.append(_rc(r) if rng.random() < 0.5 else r)
    return _readids(reads)


def _gpu(gx, text, k, steps, method=None):
    with gx.GraphBuilder(k, method=method) as b:
        b.add_reads(text)
        b.finish()
        b.run_unitigs()
        out = None
        for name, args in steps:
            out = getattr(b, name)(*args)
        return out


def _sub(s, j):
    return s[:j] + "ACGT"[("ACGT".index(s[j]) + 1) % 4] + s[j + 1:]


@pytest.mark.parametrize("k", [21, 31])
def test_alleles_match_oracle(gx, k):
    cases = [(s, lambda g: _sub(g, 300)) for s in (1, 9, 12, 17, 29)]
    cases += [(s, lambda g: g[:300] + g[301:]) for s in (1, 2, 3)]                 # deletion
    cases += [(s, lambda g: _sub(_sub(g, 300), 305)) for s in (3, 4)]              # two substitutions
    for seed, alt in cases:
        text = _alleles(seed, 600, 70, alt)
        recs = oracle.build_graph(text, k)[0]
        for maxd, maxlen in ((0.05, -1), (0.02, -1), (0.05, 2 * k - 1), (1.0, 4097)):
            got = _gpu(gx, text, k, [("clean_bubbles", (maxd, maxlen))])
            want = oracle.bubble_merge(recs, k, maxd, None if maxlen < 0 else maxlen)
            assert got == want, (seed, maxd, maxlen)


def _multi_alleles(seed, G, L, alts, fracs):
    """Reads of a genome whose window carries one of several alleles, each
    in its fraction of the reads (the rest carry the genome's)."""
    rng = np.random.default_rng(seed)
    genome = "".join("ACGT"[x] for x in rng.integers(0, 4, G))
    seqs = [genome] + [alt(genome) for alt in alts]
    cut = np.cumsum([1.0 - sum(fracs)] + list(fracs))
    reads = []
    for i, s in enumerate(range(0, G - L + 1, 2)):
        which = int(np.searchsorted(cut, ((i * 7) % 100 + 0.5) / 100.0))
        r = seqs[min(which, len(seqs) - 1)][s:s + L]
        reads.append(_rc(r) if rng.random() < 0.5 else r)
    return _readids(reads)


def _sub2(s, j):
    return s[:j] + "ACGT"[("ACGT".index(s[j]) + 2) % 4] + s[j + 1:]


@pytest.mark.parametrize("k", [21, 63])
def test_several_alleles_match_oracle(gx, k):
    """Four alleles at one position -- the genome's letter, two other
    letters and a deletion -- make one set of four bubbles between the same
    two vertices.  At thresholds between 1/(2k-1) (two substitutions apart
    by one letter) and 1/(2k-2) (the deletion) the first top takes the
    substitutions and leaves the deletion to the next round of the set."""
    a = 300
    alts = [lambda g: _sub(g, a), lambda g: _sub2(g, a), lambda g: g[:a] + g[a + 1:]]
    f1 = np.float32(1) / np.float32(2 * k - 1)
    f2 = np.float32(1) / np.float32(2 * k - 2)
    maxds = [float(np.nextafter(f1, np.float32(0))), float(f1), float((f1 + f2) / np.float32(2)), float(f2), 0.2]
    second_round = False
    for seed in (1, 2, 3):
        text = _multi_alleles(seed, 700, 2 * k + 30, alts, [0.25, 0.2, 0.15])
        recs = oracle.build_graph(text, k)[0]
        outs = []
        for maxd in maxds:
            got = _gpu(gx, text, k, [("clean_bubbles", (maxd, -1))])
            assert got == oracle.bubble_merge(recs, k, maxd), (seed, maxd)
            outs.append(got)
        second_round |= outs[2] != outs[3]  # the deletion merged only at 1/(2k-2)
    assert second_round


@pytest.mark.parametrize("k,paired", [(21, False), (31, True), (55, True), (24, True)])
def test_raw_graph_matches_oracle(gx, k, paired):
    """Sequencing errors make many bubbles in the graph before cleaning."""
    text = _sim(5100 + k, paired=paired)
    recs = oracle.build_graph(text, k)[0]
    for maxd, maxlen in ((0.05, -1), (0.2, 300)):
        got = _gpu(gx, text, k, [("clean_bubbles", (maxd, maxlen))])
        assert got == oracle.bubble_merge(recs, k, maxd, None if maxlen < 0 else maxlen), (maxd, maxlen)


@pytest.mark.parametrize("k", [31, 55])
def test_pipeline_matches_oracle(gx, k):
    """REMOVE_BAD_COVERAGE, TIP_SINGLE_NODE and BUBBLE, each with its merge
    (GenomixJobConf.java:511-517), on a diploid-like read set."""
    text = _sim(5200 + k, cov=30, err=0.005)
    recs = oracle.build_graph(text, k)[0]
    got = _gpu(gx, text, k, [("clean_bad_coverage", (3.0,)), ("clean_tips", (-1,)), ("clean_bubbles", (0.05, -1))])
    assert got == oracle.clean_pipeline(recs, k)


def test_bubbles_twice_then_coverage(gx):
    """Heads moved by a bubble merge stay where they were put through a
    second bubble merge and a coverage cleaning."""
    k = 21
    text = _sim(77, G=30000, cov=25, paired=False)
    recs = oracle.build_graph(text, k)[0]
    once = oracle.bubble_nodes(recs, k, 0.1)
    assert _gpu(gx, text, k, [("clean_bubbles", (0.1, -1)), ("clean_bubbles", (0.2, -1))]) == \
        oracle.bubble_merge(once, k, 0.2)
    assert _gpu(gx, text, k, [("clean_bubbles", (0.1, -1)), ("clean_bad_coverage", (2.0,))]) == \
        oracle.clean_bad_coverage(once, k, 2.0)


def test_hash_method(gx):
    k = 31
    text = _sim(88, G=20000)
    recs = oracle.build_graph(text, k)[0]
    assert _gpu(gx, text, k, [("clean_bubbles", (0.05, -1))], method="hash") == oracle.bubble_merge(recs, k)


def test_errors(gx):
    text = _sim(3, G=5000, cov=8)
    with gx.GraphBuilder(21) as b:
        b.add_reads(text)
        b.finish()
        with pytest.raises(gx.GxError) as e:
            b.clean_bubbles()
        assert e.value.code == gx.GX_ESTATE
        b.run_unitigs()
        plain = b.unitigs()
        for bad in (-0.1, 1.5):
            with pytest.raises(gx.GxError) as e:
                b.clean_bubbles(bad)
            assert e.value.code == gx.GX_EINVAL
        with pytest.raises(gx.GxError) as e:
            b.clean_bubbles(0.05, 5000)
        assert e.value.code == gx.GX_ENOTSUP
        assert b.clean_bubbles(math.nan) == plain  # accepted, as the reference's check lets NaN through; merges nothing


def test_handle_reused_after_bubbles(gx):
    """A build on the same handle after a bubble merge starts from the
    build's read heads again (no letters left over from the moved ones)."""
    k = 31
    text = _sim(91, G=20000)
    recs = oracle.build_graph(text, k)[0]
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        b.run_unitigs()
        assert b.clean_bubbles(0.2, -1) == oracle.bubble_merge(recs, k, 0.2)
        b.reset()
        b.add_reads(text)
        b.finish()
        assert b.part_bytes(0) == recs
        assert b.unitigs() == oracle.unitigs(recs, k)
        assert b.clean_bad_coverage(2.0) == oracle.clean_bad_coverage(recs, k, 2.0)


@pytest.mark.parametrize("text", [b"", b"1\tACGTACGTACGTTGCAACGTAG\n2\tTTGACCATGACGTAGGCAATCA\n",
                                  b"1\tACGTACGTACGTTGCAACGTACGGT\n"])
def test_graph_steps_on_empty_and_tiny_graphs(gx, text):
    """The steps after the build on a graph with no node, with two short
    reads, and with one read: the same records, statistics and FASTA as the
    oracle."""
    k = 21
    recs = oracle.build_graph(text, k)[0]
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        assert b.part_bytes(0) == recs
        assert b.unitigs() == oracle.unitigs(recs, k)
        assert b.clean_bad_coverage(1.0) == oracle.clean_bad_coverage(recs, k, 1.0)
        assert b.clean_tips(-1) == oracle.clean_coverage_then_tips(recs, k, 1.0)
        final = b.clean_bubbles(0.05, -1)
        assert final == oracle.clean_pipeline(recs, k, min_cov=1.0)
        assert b.graph_stats("unitigs") == oracle.graph_stats(final, k)
        assert b.fasta("unitigs") == oracle.dump_fasta(final, k)


def test_node_graph_after_bubbles(gx):
    """The node graph a bubble merge leaves (the kept bubble's first node
    holding the removed ones' counts, moved heads on the nodes covering
    them) has the oracle's statistics and FASTA."""
    k = 31
    text = _sim(93, G=30000, cov=25)
    recs = oracle.build_graph(text, k)[0]
    nodes = oracle.bubble_nodes(recs, k, 0.1)
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        b.run_unitigs()
        b.clean_bubbles(0.1, -1, emit=False)
        assert b.graph_stats("nodes") == oracle.graph_stats(nodes, k)
        assert b.fasta("nodes") == oracle.dump_fasta(nodes, k)
