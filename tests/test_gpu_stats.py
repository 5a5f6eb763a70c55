"""Graph statistics and FASTA dump on the device (include/genomix_stats.h)
against the oracle's restatement over the same graph's records."""

import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def _text(seed, G=30000, cov=12, paired=True):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=100, coverage=cov, error_rate=0.01, paired=paired,
                           insert_mean=300, insert_sd=25, repeat_frac=0.05, repeat_len=(300, 900), repeat_div=0.01,
                           seed=seed)
    return sim.simulate(spec)


@pytest.mark.parametrize("k,paired", [(21, False), (31, True), (55, True)])
def test_stats_and_fasta_match_oracle(gx, k, paired):
    text = _text(6100 + k, paired=paired)
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        nodes = b.part_bytes(0)
        assert b.graph_stats("nodes") == oracle.graph_stats(nodes, k)
        assert b.fasta("nodes") == oracle.dump_fasta(nodes, k)
        u = b.unitigs()
        assert b.graph_stats("unitigs") == oracle.graph_stats(u, k)
        assert b.fasta("unitigs") == oracle.dump_fasta(u, k)
        cleaned = b.clean_bad_coverage(3.0)
        assert b.graph_stats("unitigs") == oracle.graph_stats(cleaned, k)
        assert b.fasta("unitigs") == oracle.dump_fasta(cleaned, k)


def test_stats_errors(gx):
    with gx.GraphBuilder(21) as b:
        b.add_reads(_text(3, G=5000, paired=False))
        with pytest.raises(gx.GxError) as e:
            b.graph_stats("nodes")
        assert e.value.code == gx.GX_ESTATE
        b.finish()
        with pytest.raises(gx.GxError) as e:
            b.graph_stats("unitigs")
        assert e.value.code == gx.GX_ESTATE
