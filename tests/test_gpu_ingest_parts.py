"""FASTQ ingest and output partitions on the GPU, against the oracle.

FASTQ: GenomixDriver.convertAndUploadFastqToHDFS (GenomixDriver.java:665-714)
turns FASTQ into library-N.readids lines; the device conversion must give the
same text as the oracle's restatement and the same graph as building that
text.  Partitions: with P output parts every record goes to
nonneg(hashBytes(key)) % P (KmerPartitionComputerFactory.java:28-53), each
part in KmerPointable order; every part must equal the oracle's part-<p>.
"""

import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def _reads(seed, G=4000, L=90, cov=10, paired=False):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=L, coverage=cov, error_rate=0.01, paired=paired,
                           insert_mean=3 * L, insert_sd=L // 4, seed=seed)
    return sim.reads(spec)


def _gpu_fastq(gx, k, chunks1, chunks2=None, nparts=1):
    with gx.GraphBuilder(k, nparts=nparts) as b:
        if chunks2 is None:
            for c in chunks1:
                b.add_fastq(c)
        else:
            for c1, c2 in zip(chunks1, chunks2):
                b.add_fastq(c1, c2)
        b.finish()
        return b.readids(), [b.part_bytes(p) for p in range(nparts)]


@pytest.mark.parametrize("k", [21, 33])
def test_fastq_single_end(gx, k):
    from genomix_amd import sim
    m1, _ = _reads(31 + k)
    fq = sim.fastq_text(m1)
    rid, parts = _gpu_fastq(gx, k, [fq])
    want = oracle.fastq_to_readids(fq)
    assert rid == want
    assert parts[0] == oracle.build_graph(want, k)[0]


@pytest.mark.parametrize("k", [21, 55])
def test_fastq_paired_end_in_chunks(gx, k):
    from genomix_amd import sim
    m1, m2 = _reads(77 + k, paired=True)
    f1, f2 = sim.fastq_text(m1, "a"), sim.fastq_text(m2, "b")
    # chunk boundaries fall mid-line and differ between the two files
    c1 = [f1[:1001], f1[1001:5003], f1[5003:]]
    c2 = [f2[:17], f2[17:9999], f2[9999:]]
    rid, parts = _gpu_fastq(gx, k, c1, c2)
    want = oracle.fastq_to_readids(f1, f2)
    assert rid == want
    assert parts[0] == oracle.build_graph(want, k)[0]


def test_fastq_line_edge_cases(gx):
    f1 = (b"@r0\r\n  ACGTACGTAA \r\n+\r\nIIIIIIIIII\r\n"   # CRLF, spaces around the sequence
          b"@r1\rGGCCTTAAGG\t\r+\rIIIIIIIIII\r"           # lone CR, trailing tab
          b"@r2\nacgtNNACGT\n+\nIIIIIIIIII\n"             # lowercase / N kept as is
          b"@r3\nGATTACA\n+\nIIIIIII\n"
          b"@r4\nTTTTGGGGCCAA")                            # truncated record, no final newline
    f2 = (b"@r0\nCCCCAAAATT\n+\nIIIIIIIIII\n"
          b"@r1\n\x0bGATTACAGAT\n+\nIIIIIIIIII\n"
          b"@r2\nAAAACCCCGG\n+\nIIIIIIIIII\n"
          b"@r3\nACGTTGCA\n+\nIIIIIIII\n"
          b"@r4\nCATCATCATC\n")
    for k in (3, 5):
        rid, parts = _gpu_fastq(gx, k, [f1], [f2])
        want = oracle.fastq_to_readids(f1, f2)
        assert rid == want
        assert parts[0] == oracle.build_graph(want, k)[0]
    rid, parts = _gpu_fastq(gx, 5, [f1])
    want = oracle.fastq_to_readids(f1)
    assert rid == want
    assert parts[0] == oracle.build_graph(want, 5)[0]


def test_fastq_empty_input(gx):
    rid, parts = _gpu_fastq(gx, 21, [b""])
    assert rid == b"" and parts[0] == b""


def test_fastq_empty_sequence_line_is_a_parse_error(gx):
    # "14\t" has one field after split("\t"): the parser rejects the line
    f1 = b"@r0\nACGTACGT\n+\nIIIIIIII\n@r1\n  \n+\n\n"
    with pytest.raises(oracle.OracleError):
        oracle.build_graph(oracle.fastq_to_readids(f1), 5)
    with pytest.raises(gx.GxError) as e:
        _gpu_fastq(gx, 5, [f1])
    assert e.value.code == gx.GX_EINPUT


def test_fastq_line_count_mismatch(gx):
    f1 = b"@r0\nACGTACGT\n+\nIIIIIIII\n@r1\nACGTACGT\n+\nIIIIIIII\n"
    f2 = b"@r0\nACGTACGT\n+\nIIIIIIII\n"
    with pytest.raises(oracle.OracleError):
        oracle.fastq_to_readids(f1, f2)
    with pytest.raises(gx.GxError) as e:
        _gpu_fastq(gx, 5, [f1], [f2])
    assert e.value.code == gx.GX_EINPUT


def test_input_formats_do_not_mix(gx):
    with gx.GraphBuilder(5) as b:
        b.add_reads(b"2\tACGTACGT\n")
        with pytest.raises(gx.GxError) as e:
            b.add_fastq(b"@r\nACGT\n+\nIIII\n")
        assert e.value.code == gx.GX_EINVAL
        b.finish()
        with pytest.raises(gx.GxError) as e:
            b.readids()
        assert e.value.code == gx.GX_ESTATE


@pytest.mark.parametrize("k,P", [(21, 2), (21, 7), (33, 3), (55, 64), (4, 5)])
def test_partitions_match_oracle(gx, k, P):
    from genomix_amd import sim
    m1, m2 = _reads(500 + P, paired=True)
    text = sim.readids_text(m1, m2)
    got = gx.build_graph(text, k, nparts=P)
    want = oracle.build_graph(text, k, nparts=P)
    assert len(got) == P
    for p in range(P):
        assert got[p] == want[p], f"part {p} differs"
    # the parts together hold the P = 1 records
    one = oracle.build_graph(text, k)[0]
    assert sum(len(x) for x in got) == len(one)


def test_partition_info_and_bounds(gx):
    text = b"2\tACGTTGCAAC\n6\tTTGACCAGTA\n"
    with gx.GraphBuilder(3, nparts=4) as b:
        b.add_reads(text)
        b.finish()
        want = oracle.build_graph(text, 3, nparts=4)
        for p in range(4):
            n, nb = b.part_info(p)
            assert nb == len(want[p])
            assert n == len(oracle.dump_text(want[p]).splitlines())
        for bad in (-1, 4):
            with pytest.raises(gx.GxError) as e:
                b.part_info(bad)
            assert e.value.code == gx.GX_EINVAL


def test_unitigs_do_not_depend_on_partitions(gx):
    from genomix_amd import sim
    m1, m2 = _reads(909, paired=True)
    text = sim.readids_text(m1, m2)
    outs = []
    for P in (1, 6):
        with gx.GraphBuilder(31, nparts=P) as b:
            b.add_reads(text)
            b.finish()
            outs.append(b.unitigs())
    assert outs[0] == outs[1]
    assert outs[0] == oracle.unitigs(oracle.build_graph(text, 31)[0], 31)
