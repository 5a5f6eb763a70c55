"""The oracle's BUBBLE + MERGE (oracle/genomix_oracle.cpp bubble_merge,
after SimpleBubbleMergeVertex.java:51-320).  The reference's bubble
fixtures (genomix-hadoop graphs/bubblemerge) are inputs in the retired text
format without expected outputs, so closed forms pin the restatement: a
heterozygous SNP read at 30% collapses into the majority allele and gives
exactly the merged graph of the same reads with the minority allele
corrected -- the genome as one vertex, its coverage, and every read head at
the offset the corrected read has (heads moved from the other allele keep
their offsets, the two paths being equally long); the same for a one-letter
deletion, up to the scaled offsets of the moved heads; and the
dissimilarity and length thresholds at their boundaries."""

import numpy as np
import pytest

from genomix_amd import formats
from oracle import oracle


def _readids(seqs):
    return b"".join(b"%d\t%s\n" % (4 * i + 2, s.encode()) for i, s in enumerate(seqs))


def _rc(s):
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


def _sub(s, j):
    return s[:j] + "ACGT"[("ACGT".index(s[j]) + 1) % 4] + s[j + 1:]


def _heads(hs):
    return sorted((h.offset, h.mate_id, h.read_id) for h in hs or [])


def _allele_reads(seed, G, L, k, alt_of, frac=0.3, step=2):
    """Reads tiling a random genome (half reverse-complemented); a fraction
    carries the alternative allele alt_of(genome).  Returns the reads, the
    same reads corrected to the genome, and the genome."""
    rng = np.random.default_rng(seed)
    genome = "".join("ACGT"[x] for x in rng.integers(0, 4, G))
    alt = alt_of(genome)
    reads, fixed = [], []
    for i, s in enumerate(range(0, G - L + 1, step)):
        use_alt = (i % 10) < int(frac * 10)
        r = (alt if use_alt else genome)[s:s + L]
        f = genome[s:s + L]
        if rng.random() < 0.5:
            r, f = _rc(r), _rc(f)
        reads.append(r)
        fixed.append(f)
    return reads, fixed, genome


def _same_allele_orientation(recs, k, genome, snp):
    """Whether the two allele vertices of the merged graph are stored in the
    same orientation (their sequences differ in one letter as stored)."""
    m = formats.decode_records(oracle.unitigs(recs, k))
    short = [n.internal or key for key, n in m if len(n.internal or key) == 2 * k - 1]
    assert len(short) == 2
    a, b = short
    return sum(x != y for x, y in zip(a, b)) == 1


@pytest.mark.parametrize("k,seeds", [(21, (1, 2, 3, 9, 29)), (31, (1, 2, 12, 17))])
def test_snp_collapses_to_the_corrected_graph(k, seeds):
    G, L, snp = 600, 70, 300
    orient = set()
    for seed in seeds:
        reads, fixed, genome = _allele_reads(seed, G, L, k, lambda g: _sub(g, snp))
        recs = oracle.build_graph(_readids(reads), k)[0]
        assert len(formats.decode_records(oracle.unitigs(recs, k))) == 4  # two flanks, two alleles
        orient.add(_same_allele_orientation(recs, k, genome, snp))
        got = formats.decode_records(oracle.bubble_merge(recs, k))
        want = formats.decode_records(oracle.unitigs(oracle.build_graph(_readids(fixed), k)[0], k))
        assert len(got) == len(want) == 1
        (gk, g), (wk, w) = got[0], want[0]
        assert gk == wk and g.internal == w.internal and g.internal in (genome, _rc(genome))
        assert g.coverage == w.coverage
        assert g.edges == w.edges
        assert _heads(g.unflipped) == _heads(w.unflipped) and _heads(g.flipped) == _heads(w.flipped)
    # both stored orientations of the two alleles occur among the seeds
    assert orient == {True, False}


def test_indel_collapses_with_scaled_heads():
    """A one-letter deletion: alleles of 2k-1 and 2k-2 letters, edit
    distance 1.  The majority allele stays; moved heads land at offsets
    (o + 1) f - f, f = the length ratio, so all stay within the genome and
    keep their count and ids."""
    k, G, L, pos = 21, 600, 70, 300
    for seed in range(1, 5):
        reads, fixed, genome = _allele_reads(seed, G, L, k, lambda g: g[:pos] + g[pos + 1:])
        recs = oracle.build_graph(_readids(reads), k)[0]
        got = formats.decode_records(oracle.bubble_merge(recs, k))
        want = formats.decode_records(oracle.unitigs(oracle.build_graph(_readids(fixed), k)[0], k))
        assert len(got) == 1
        g, w = got[0][1], want[0][1]
        assert g.internal == w.internal and g.coverage == pytest.approx(w.coverage, rel=1e-6)
        ids = lambda n: sorted((h.mate_id, h.read_id) for h in (n.unflipped or []) + (n.flipped or []))
        assert ids(g) == ids(w)
        assert all(0 <= h.offset < G for h in (g.unflipped or []) + (g.flipped or []))
        _moved_heads_keep_their_relative_place(recs, k, g)


def _moved_heads_keep_their_relative_place(recs, k, merged):
    """Heads of the removed allele sit where they were relative to its
    length: at o' * Lt / Lo letters into the kept allele (within one
    letter, the reference truncating), o' counted from the end the two
    alleles share."""
    m = formats.decode_records(oracle.unitigs(recs, k))
    alleles = sorted((n for _, n in m if n.internal and len(n.internal) < 2 * k), key=lambda n: -n.coverage)
    top, bub = alleles
    Lt, Lo = len(top.internal), len(bub.internal)
    assert Lt != Lo
    same = _ed(top.internal, bub.internal) <= _ed(top.internal, _rc(bub.internal))
    seq = merged.internal
    start = seq.find(top.internal)
    fwd = start >= 0
    if not fwd:
        start = seq.find(_rc(top.internal))
    assert start >= 0
    final = {h.read_id: h.offset for h in (merged.unflipped or []) + (merged.flipped or [])}
    moved = (bub.unflipped or []) + (bub.flipped or [])
    assert moved
    for h in moved:
        o = h.offset if same else Lo - 1 - h.offset
        y = final[h.read_id]
        new = y - start if fwd else start + Lt - 1 - y
        assert o * Lt / Lo - 1 < new <= o * Lt / Lo + 1e-9, (o, new)


def _ed(a, b):
    prev = list(range(len(b) + 1))
    for i, x in enumerate(a, 1):
        cur = [i]
        for j, y in enumerate(b, 1):
            cur.append(min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (x != y)))
        prev = cur
    return prev[-1]


def test_dissimilarity_threshold():
    """Two substitutions 5 letters apart: one bubble of 2k+4 letters per
    allele at edit distance 2; merged iff 2 / (2k + 4) <= maxDissimilarity."""
    k, G, L, a = 21, 600, 80, 300
    reads, fixed, genome = _allele_reads(3, G, L, k, lambda g: _sub(_sub(g, a), a + 5))
    recs = oracle.build_graph(_readids(reads), k)[0]
    plain = oracle.unitigs(recs, k)
    frac = np.float32(2) / np.float32(2 * k + 4)
    assert oracle.bubble_merge(recs, k, float(frac)) != plain
    assert len(formats.decode_records(oracle.bubble_merge(recs, k, float(frac)))) == 1
    below = float(np.nextafter(frac, np.float32(0)))
    assert oracle.bubble_merge(recs, k, below) == plain
    assert oracle.bubble_merge(recs, k) == oracle.bubble_merge(recs, k, float(frac))  # default 0.05 > 2/46


def test_length_threshold():
    """A bubble of 2k - 1 letters is one when maxLength exceeds that."""
    k, G, L = 21, 600, 70
    reads, _, _ = _allele_reads(4, G, L, k, lambda g: _sub(g, 300))
    recs = oracle.build_graph(_readids(reads), k)[0]
    plain = oracle.unitigs(recs, k)
    assert oracle.bubble_merge(recs, k, 0.05, 2 * k - 1) == plain
    assert len(formats.decode_records(oracle.bubble_merge(recs, k, 0.05, 2 * k))) == 1
    assert oracle.bubble_merge(recs, k) == oracle.bubble_merge(recs, k, 0.05, 5 * k)  # the default


def test_no_bubble_keeps_the_merge():
    rng = np.random.default_rng(9)
    genome = "".join("ACGT"[x] for x in rng.integers(0, 4, 500))
    reads = [genome[s:s + 60] for s in range(0, 441, 3)]
    recs = oracle.build_graph(_readids(reads), 21)[0]
    assert oracle.bubble_merge(recs, 21) == oracle.unitigs(recs, 21)
