"""Key-range passes and the record sink (include/genomix_stream.h): a build
taken in several ranges of the sort key's top byte gives, partition by
partition, the bytes of a single-pass build and of the oracle; a sink
receives the same bytes range by range."""

import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def _text(seed, G=40000, L=100, cov=20, paired=True):
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=G, read_len=L, coverage=cov, error_rate=0.01, paired=paired,
                           insert_mean=3 * L, insert_sd=L // 4, repeat_frac=0.05, repeat_len=(300, 900),
                           repeat_div=0.01, seed=seed)
    return sim.simulate(spec)


def _parts(b, P):
    return [b.part_bytes(p) for p in range(P)]


@pytest.mark.parametrize("k,P", [(21, 1), (31, 3), (55, 2), (33, 1)])
def test_ranges_match_single_pass_and_oracle(gx, k, P):
    """Every partition's bytes, the statistics and, from the node graph the
    ranges keep in HBM, the unitig records, cleaning and GraphStatistics
    equal a single-pass build's (and the oracle's)."""
    text = _text(900 + k, paired=k != 33)
    with gx.GraphBuilder(k, nparts=P) as b:
        b.add_reads(text)
        b.finish()
        want = _parts(b, P)
        info1 = [b.part_info(p) for p in range(P)]
        st1 = b.stats()
        assert st1["key_ranges"] == 1
        nstats1 = b.graph_stats("nodes")
        u1 = b.unitigs()
        ustats1 = b.graph_stats("unitigs")
        c1 = b.clean_bad_coverage(3.0)
    want_all = oracle.build_graph(text, k, nparts=P)
    assert want == want_all
    one_part = want_all[0] if P == 1 else oracle.build_graph(text, k)[0]
    assert u1 == oracle.unitigs(one_part, k)
    assert c1 == oracle.clean_bad_coverage(one_part, k, 3.0)
    N = st1["occurrences"]
    for per in (N // 5, N // 40, 1):
        with gx.GraphBuilder(k, nparts=P) as b:
            b.set_stream(range_occurrences=per)
            b.add_reads(text)
            b.finish()
            st = b.stats()
            assert st["key_ranges"] > 1
            if per == 1:
                assert st["key_ranges"] >= 200  # one range per top-byte value that occurs
            assert _parts(b, P) == want, f"range_occurrences={per}"
            for key in ("occurrences", "distinct", "heads", "replayed_nodes", "replayed_occurrences"):
                assert st[key] == st1[key], key
            assert [b.part_info(p) for p in range(P)] == info1
            assert b.graph_stats("nodes") == nstats1, f"node statistics, range_occurrences={per}"
            assert b.unitigs() == u1, f"unitigs, range_occurrences={per}"
            assert b.graph_stats("unitigs") == ustats1
            assert b.clean_bad_coverage(3.0) == c1
    with gx.GraphBuilder(k) as b:  # a one-range build keeps the unitig pass too
        b.set_stream(range_occurrences=10 * N)
        b.add_reads(text)
        b.finish()
        assert b.stats()["key_ranges"] == 1
        assert b.unitigs() == u1


def test_bubbles_and_fastq_after_ranges(gx):
    """BUBBLE on the graph a key-range build kept equals BUBBLE after a
    one-pass build and the oracle's pipeline; paired FASTQ converted on the
    device and built in ranges gives the one-pass bytes and unitigs."""
    from genomix_amd import sim
    k = 31
    text = _text(2468, G=30000, cov=25)
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        b.run_unitigs()
        want_bub = b.clean_bubbles(0.1)
        N = b.stats()["occurrences"]
    assert want_bub == oracle.bubble_merge(oracle.build_graph(text, k)[0], k, 0.1)
    with gx.GraphBuilder(k) as b:
        b.set_stream(range_occurrences=N // 7)
        b.add_reads(text)
        b.finish()
        assert b.stats()["key_ranges"] > 1
        b.run_unitigs()
        assert b.clean_bubbles(0.1) == want_bub
    spec = sim.ReadSetSpec(genome_len=20000, read_len=100, coverage=15, error_rate=0.01, paired=True,
                           insert_mean=300, insert_sd=25, seed=1357)
    m1, m2 = sim.reads(spec)
    f1, f2 = sim.fastq_text(m1), sim.fastq_text(m2)
    got = []
    for per in (0, 20_000):
        with gx.GraphBuilder(k) as b:
            b.set_stream(range_occurrences=per)
            b.add_fastq(f1, f2)
            b.finish()
            assert (b.stats()["key_ranges"] > 1) == (per > 0)
            got.append((b.readids(), b.part_bytes(0), b.unitigs()))
    assert got[0] == got[1]
    assert got[0][1] == oracle.build_graph(oracle.fastq_to_readids(f1, f2), k)[0]


def test_ranges_twice_on_one_handle(gx):
    """A second key-range build on the same handle (its buffers, and the
    node graph of the first, reused or reclaimed) gives the same unitigs;
    a smaller read set after it too."""
    k = 31
    big, small = _text(4242), _text(4343, G=15000)
    want = {}
    for name, text in (("big", big), ("small", small)):
        with gx.GraphBuilder(k) as b:
            b.add_reads(text)
            b.finish()
            want[name] = (b.part_bytes(0), b.unitigs())
    with gx.GraphBuilder(k) as b:
        for name, text in (("big", big), ("big", big), ("small", small), ("big", big)):
            b.set_stream(range_occurrences=len(text) // 20)
            b.reset()
            b.add_reads(text)
            b.finish()
            assert b.stats()["key_ranges"] > 1
            assert (b.part_bytes(0), b.unitigs()) == want[name], name


@pytest.mark.parametrize("per", [0, 1 << 15])
def test_repeated_builds_reuse_every_buffer(gx, per):
    """A build repeated on its handle allocates nothing once the handle's
    buffers have settled (one pass, and key ranges that keep their graph),
    and gives the same bytes and unitigs."""
    text = _text(31337)
    with gx.GraphBuilder(31) as b:
        b.set_stream(range_occurrences=per)
        out = []
        for step in range(4):
            b.reset()
            b.add_reads(text)
            b.finish()
            st = b.stats()
            out.append((b.part_bytes(0), b.unitigs()))
            if per:
                assert st["key_ranges"] > 1
            if step >= 2:
                assert st["device_allocations"] == 0, f"step {step}"
                assert st["buffers_reclaimed"] == 0, f"step {step}"
        assert all(o == out[0] for o in out)


@pytest.mark.parametrize("method", ["sort", "hash"])
def test_sink_receives_each_partition_in_order(gx, method):
    from genomix_amd import devmem
    k, P = 31, 4
    text = _text(77)
    want = oracle.build_graph(text, k, nparts=P)
    got = [[] for _ in range(P)]
    nrec = [0] * P

    def sink(part, addr, nbytes, n):
        got[part].append(devmem.to_host(addr, nbytes))
        nrec[part] += n
        return 0

    with gx.GraphBuilder(k, nparts=P, method=method) as b:
        b.set_stream(range_occurrences=50_000 if method == "sort" else 0, sink=sink)
        b.add_reads(text)
        b.finish()
        st = b.stats()
        assert st["key_ranges"] > 1 if method == "sort" else st["key_ranges"] == 1
        assert [b"".join(c) for c in got] == want
        for p in range(P):
            assert b.part_info(p) == (nrec[p], len(want[p]))
        with pytest.raises(gx.GxError) as e:
            b.part_bytes(0)
        assert e.value.code == gx.GX_ESTATE
        if method == "sort":  # the graph left range by range: no unitig pass over it
            with pytest.raises(gx.GxError) as e:
                b.run_unitigs()
            assert e.value.code == gx.GX_ENOTSUP
        # back to records kept in HBM on the same handle
        b.set_stream()
        b.reset()
        b.add_reads(text)
        b.finish()
        assert _parts(b, P) == want


def test_sink_error_aborts_the_build(gx):
    text = _text(5, G=10000)

    def bad(part, addr, nbytes, n):
        raise ValueError("consumer full")

    def code(part, addr, nbytes, n):
        return -12

    with gx.GraphBuilder(21) as b:
        b.set_stream(range_occurrences=1000, sink=bad)
        b.add_reads(text)
        with pytest.raises(ValueError, match="consumer full"):
            b.finish()
    with gx.GraphBuilder(21) as b:
        b.set_stream(sink=code)
        b.add_reads(text)
        with pytest.raises(gx.GxError) as e:
            b.finish()
        assert e.value.code == -12


@pytest.mark.parametrize("k", [31, 55])
def test_large_pass_three_byte_buckets(gx, k):
    """A pass of more than 2^28 records sorts in buckets of the top three key
    bytes; the same read set in smaller passes (two-byte buckets) and through
    the hash-table aggregation gives the same bytes, and the same unitigs
    (after the passes: from the node graph they kept, with a node index of
    more than 2^26 nodes built over all of them)."""
    from genomix_amd import sim
    spec = sim.ReadSetSpec(genome_len=6_000_000, read_len=150, coverage=70, error_rate=0.01, paired=True,
                           insert_mean=500, insert_sd=50, repeat_frac=0.05, repeat_len=2000, repeat_div=0.01,
                           seed=20261299)
    text = sim.simulate(spec)
    with gx.GraphBuilder(k) as b:
        b.add_reads(text)
        b.finish()
        st = b.stats()
        # more than 2^28 records in one pass, more than 2^26 nodes (three-byte
        # node index directory)
        assert st["occurrences"] > 1 << 28 and st["key_ranges"] == 1 and st["distinct"] > 1 << 26
        one = b.part_bytes(0)
        b.run_unitigs()  # every neighbour lookup goes through the node index
        nu = b.unitig_count()
        assert 0 < nu < st["distinct"]
        uone = b.unitig_records()
        b.set_stream(range_occurrences=st["occurrences"] // 3)
        b.reset()
        b.add_reads(text)
        b.finish()
        assert b.stats()["key_ranges"] >= 3
        assert b.part_bytes(0) == one
        assert b.unitigs() == uone
    with gx.GraphBuilder(k, method="hash") as b:
        b.add_reads(text)
        b.finish()
        assert b.part_bytes(0) == one
        b.run_unitigs()
        assert b.unitig_count() == nu
