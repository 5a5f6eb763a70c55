"""The oracle's REMOVE_BAD_COVERAGE + MERGE (oracle/genomix_oracle.cpp
clean_bad_coverage, after RemoveBadCoverageVertex.java:38-70).  The
reference's only fixture for it (TestSet LOW_COVERAGE/MaxCoverage_6) is in
the retired text format and its thresholds are not in the tree, so these
checks pin the restatement instead: bounds that keep everything give the
plain merge; every merged record left lies within the bounds; and a unique
genome read with one sequencing error cleans back to exactly the merged
graph of the error-free reads (one unitig spelling the genome)."""

import math

import numpy as np

from genomix_amd import formats, sim
from oracle import oracle


def _reads(seed, G=4000, cov=10):
    spec = sim.ReadSetSpec(genome_len=G, read_len=80, coverage=cov, error_rate=0.01, seed=seed)
    return sim.simulate(spec)


def test_bounds_keeping_everything_give_the_plain_merge():
    for seed, k in ((1, 21), (2, 31), (3, 9)):
        recs = oracle.build_graph(_reads(seed), k)[0]
        assert oracle.clean_bad_coverage(recs, k, 0.0, math.inf) == oracle.unitigs(recs, k)


def test_records_left_lie_within_the_bounds():
    recs = oracle.build_graph(_reads(5, cov=15), 21)[0]
    for lo, hi in ((3.0, math.inf), (5.0, 40.0), (1.5, 9.0)):
        out = formats.decode_records(oracle.clean_bad_coverage(recs, 21, lo, hi))
        assert out, (lo, hi)
        assert all(lo <= n.coverage <= hi for _, n in out)
    assert oracle.clean_bad_coverage(recs, 21, 10.0, 5.0) == b""  # nothing can satisfy both


def _readids(seqs):
    return b"".join(b"%d\t%s\n" % (4 * i + 2, s.encode()) for i, s in enumerate(seqs))


def _rc(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def test_one_error_cleans_back_to_the_genome():
    rng = np.random.default_rng(11)
    G, L, k = 600, 60, 21
    genome = "".join("ACGT"[x] for x in rng.integers(0, 4, G))
    starts = list(range(0, G - L + 1, 3))  # every base covered by ~20 reads
    reads = [genome[s:s + L] for s in starts]
    clean = oracle.unitigs(oracle.build_graph(_readids(reads), k)[0], k)
    seqs = [n.internal for _, n in formats.decode_records(clean)]
    assert len(seqs) == 1 and seqs[0] in (genome, _rc(genome))
    bad = list(reads)
    mid = bad[40]
    bad[40] = mid[:30] + "ACGT"[("ACGT".index(mid[30]) + 1) % 4] + mid[31:]  # one substitution
    recs = oracle.build_graph(_readids(bad), k)[0]
    assert len(formats.decode_records(oracle.unitigs(recs, k))) > 1  # the error splits the genome path
    cleaned = formats.decode_records(oracle.clean_bad_coverage(recs, k, 3.0))
    assert [n.internal for _, n in cleaned] == seqs


def test_end_error_tip_is_removed():
    """An error on a read's last base makes one erroneous k-mer hanging off
    the genome path (a single-node tip); TIP_SINGLE_NODE removes it and the
    merge gives back the genome as one unitig.  An error mid-read makes a
    longer branch, which the rule keeps."""
    rng = np.random.default_rng(12)
    G, L, k = 600, 60, 21
    genome = "".join("ACGT"[x] for x in rng.integers(0, 4, G))
    reads = [genome[s:s + L] for s in range(0, G - L + 1, 3)]
    base = formats.decode_records(oracle.unitigs(oracle.build_graph(_readids(reads), k)[0], k))
    assert len(base) == 1
    sub = lambda s, j: s[:j] + "ACGT"[("ACGT".index(s[j]) + 1) % 4] + s[j + 1:]
    tip = list(reads)
    tip[40] = sub(tip[40], L - 1)
    recs = oracle.build_graph(_readids(tip), k)[0]
    assert len(formats.decode_records(oracle.unitigs(recs, k))) == 3  # the genome in two at the branch, and the tip
    cleaned = formats.decode_records(oracle.clean_tips(recs, k))
    assert [n.internal for _, n in cleaned] == [base[0][1].internal]
    mid = list(reads)
    mid[40] = sub(mid[40], 30)
    recs = oracle.build_graph(_readids(mid), k)[0]
    assert oracle.clean_tips(recs, k) == oracle.unitigs(recs, k)
