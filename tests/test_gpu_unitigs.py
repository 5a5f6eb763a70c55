"""Parity of the HIP unitig compression with the oracle's sequential walker
(byte for byte: both fix P4's random survivor and orientation the same way),
plus the RandMergeOffsetTest invariant on the device path."""

import glob
import os
import random
import re

import pytest

from oracle import oracle

from .conftest import GOLDEN
from .test_oracle_unitigs import _genome_without_repeats, field, lines_of, rc, tiled_reads

pytestmark = pytest.mark.gpu

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "hyracks", "**", "*.test.txt"), recursive=True))


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def check(gx, text, k):
    g_gpu, u_gpu = gx.build_and_merge(text, k)
    g_ora = oracle.build_graph(text, k)[0]
    assert g_gpu == g_ora
    u_ora = oracle.unitigs(g_ora, k)
    if u_gpu != u_ora:
        a, b = lines_of(u_gpu), lines_of(u_ora)
        for x, y in zip(a, b):
            if x != y:
                pytest.fail(f"first differing unitig:\n gpu: {x}\n ora: {y}")
        pytest.fail(f"unitig counts {len(a)} vs {len(b)}")


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def test_fixture_unitigs(gx, path):
    k = int(re.search(r"_k(\d+)\.", os.path.basename(path)).group(1))
    check(gx, open(path, "rb").read(), k)


@pytest.mark.parametrize("k", [3, 4, 5, 11, 21, 31, 32, 33, 55, 63])
def test_random_unitigs(gx, k):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=4000, read_len=k + 60, coverage=15, error_rate=0.01, paired=True,
                           insert_mean=3 * (k + 60), insert_sd=20, repeat_frac=0.1, repeat_len=300,
                           repeat_div=0.02, seed=500 + k)
    check(gx, sim.simulate(spec), k)


@pytest.mark.parametrize("k", [3, 4, 5, 6, 7, 8, 9, 11, 12])
def test_overlap_graph_unitigs(gx, k):
    """Graphs holding every (k-1)-overlap of their k-mers: all four outer
    edge types, self edges of cycles, palindromic neighbours at even k; the
    oracle's records here are pinned by sequence-only brute force."""
    from .test_oracle_merge_pins import overlap_graph_text, random_kmer_set
    for seed in range(6):
        text, _ = overlap_graph_text(random_kmer_set(random.Random(1000 * k + seed), k), k)
        check(gx, text, k)


@pytest.mark.parametrize("k,G,cov", [(21, 30000, 20), (31, 60000, 40), (33, 20000, 30), (55, 40000, 30)])
def test_walk_ranking_equals_pointer_jumping(gx, k, G, cov):
    """Odd k ranks by walks; forcing pointer jumping gives the same records,
    and both equal the oracle's."""
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=150, coverage=cov, error_rate=0.01, paired=True,
                           insert_mean=500, insert_sd=50, repeat_frac=0.05, repeat_len=2000, repeat_div=0.01,
                           seed=900 + k)
    text = sim.simulate(spec)
    outs, rounds = [], []
    for mode in ("auto", "pointer_jumping"):
        with gx.GraphBuilder(k) as b:
            b.set_ranking(mode)
            b.add_reads(text)
            b.finish()
            outs.append(b.unitigs())
            rounds.append(b.stats()["pj_rounds"])
    assert rounds[0] == 0 and rounds[1] > 0  # the walks ranked every node; the forced mode jumped
    assert outs[0] == outs[1]
    assert outs[0] == oracle.unitigs(oracle.build_graph(text, k)[0], k)


def test_walk_falls_back_on_cycles_and_long_paths(gx):
    """A circular genome (a cycle) and a path longer than the two walkers'
    caps together are ranked by pointer jumping in the default mode, with the
    oracle's records."""
    rng = random.Random(77)
    k = 21
    genome = "".join(rng.choice("ACGT") for _ in range(6000))
    circ = genome + genome[:200]
    reads = [circ[p:p + 120] for p in range(0, 6000, 40)]
    text = "".join(f"{4 * i + 2}\t{s}\n" for i, s in enumerate(reads)).encode()
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        got = b.unitigs()
        assert b.stats()["pj_rounds"] > 0
    assert got == oracle.unitigs(oracle.build_graph(text, k)[0], k)
    # the two walkers of a path each claim at most 4096 nodes: a path of
    # ~9900 nodes leaves nodes unclaimed, whichever walker runs ahead
    long_genome = genome + "".join(rng.choice("ACGT") for _ in range(4000))
    lin = "".join(f"{4 * i + 2}\t{long_genome[p:p + 120]}\n"
                  for i, p in enumerate(range(0, len(long_genome) - 120, 40))).encode()
    with gx.GraphBuilder(k) as b:
        b.add_reads(lin)
        b.finish()
        got = b.unitigs()
        assert b.stats()["pj_rounds"] > 0
    assert got == oracle.unitigs(oracle.build_graph(lin, k)[0], k)


def test_cycles_and_tandems(gx):
    rng = random.Random(3)
    lines = []
    # circular genomes (reads wrap around) make cycles; short units make tandems
    for c in range(20):
        G = rng.randint(12, 40)
        g = "".join(rng.choice("ACGT") for _ in range(G))
        gg = g * 4
        for j in range(8):
            p = rng.randrange(G)
            s = gg[p:p + rng.randint(10, 30)]
            lines.append(s if rng.random() < 0.5 else rc(s))
    for u in ("A", "AC", "ACG", "AAT"):
        lines.append(u * 12)
    text = "".join(f"{4 * i + 2}\t{s}\n" for i, s in enumerate(lines)).encode()
    for k in (5, 6, 7, 9):
        check(gx, text, k)


@pytest.mark.parametrize("k,G", [(21, 1200), (31, 2000), (55, 2000)])
def test_single_genome_one_unitig_on_gpu(gx, k, G):
    rng = random.Random(k)
    genome = _genome_without_repeats(rng, G, k)
    reads = tiled_reads(genome, 100, 7, rng)
    text = "".join(f"{4 * i + 2}\t{s}\n" for i, (_, _, s) in enumerate(reads)).encode()
    _, u = gx.build_and_merge(text, k)
    ls = lines_of(u)
    assert len(ls) == 1
    assert field(ls[0], "kmer") in (genome, rc(genome))


def test_plain_gpu_workload_unitig_invariants(gx):
    """5 Mbp workload at full size: every node lands in exactly one unitig
    (sum of unitig lengths = distinct k-mers), the k-mer counts add up
    (sum of coverage x length = occurrences, within float rounding), and the
    pass is deterministic."""
    from genomix_amd import sim
    text = sim.simulate(sim.PLAIN_GPU_5MBP)
    with gx.GraphBuilder(31) as b:
        b.add_reads(text)
        b.finish()
        st = b.stats()
        u = b.unitigs()
        assert b.unitigs() == u
    nrec, total_len, total_cov, ascending = oracle.unitig_summary(u, 31)
    assert ascending
    assert nrec < st["distinct"]
    assert total_len == st["distinct"]
    assert abs(total_cov - st["occurrences"]) < 1e-5 * st["occurrences"]


def test_head_beyond_offset_field_is_refused(gx):
    """One error-free unitig longer than 2^23 letters: the heads of the reads
    near its far end would need offsets beyond ReadHeadInfo's 23 bits, where
    the reference's makeUUID throws (ReadHeadInfo.java:115-119); the unitig
    pass fails with GX_ENOTSUP instead of writing truncated offsets."""
    import numpy as np
    from genomix_amd import sim
    lim = gx.unitig_limits()["max_head_offset"]
    G, L, step = lim + 200_000, 150, 100  # tiled reads: every (k+1)-mer is read
    rng = np.random.default_rng(23)
    g = rng.integers(0, 4, G, dtype=np.uint8)
    starts = np.arange(0, G - L + 1, step)
    starts[-1] = G - L
    idx = starts[:, None] + np.arange(L)[None, :]
    reads = g[idx]
    flip = rng.random(len(starts)) < 0.5
    reads[flip] = (3 - reads[flip])[:, ::-1]
    text = sim.readids_text_np(reads)
    with gx.GraphBuilder(31) as b:
        b.add_reads(text)
        b.finish()
        with pytest.raises(gx.GxError) as e:
            b.run_unitigs()
        assert e.value.code == gx.GX_ENOTSUP
        assert "23 bits" in str(e.value)
    with pytest.raises(oracle.OracleError):
        oracle.make_uuid(0, 0, 2, lim + 1)
