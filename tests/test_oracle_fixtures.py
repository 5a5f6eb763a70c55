"""The CPU oracle against the reference's own fixtures.

Mirrors genomix-hyracks BatchTestFilesTest (BatchTestFilesTest.java:48-172):
each ``*_k<K>.test.txt`` readids input is built and its text dump compared
with ``*.expected.txt`` after sorting lines, edge lists (fields 1-4) as sets
(TestUtils.compareFilesWithUnOrderedFields, TestUtils.java:67-176) and every
other field exactly.  Beyond that, the edge-list ORDER of every node is
replayed through the reference's split/partition structure.
"""

import glob
import itertools
import os
import re

import pytest

from oracle import oracle

from .conftest import GOLDEN

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "hyracks", "**", "*.expected.txt"), recursive=True))


def _k_of(path):
    return int(re.search(r"_k(\d+)\.", os.path.basename(path)).group(1))


def _unordered(line):
    f = line.split("\t")
    for i in (1, 2, 3, 4):
        f[i] = "".join(sorted(re.split("[^ATCG]+", f[i])))
    return "\t".join(f)


def _lists(line):
    f = line.split("\t")
    out = []
    for i in (1, 2, 3, 4):
        v = f[i].split(":", 1)[1]
        out.append([] if v == "null" else v[1:-1].split(","))
    return f[0], out


def hadoop_splits(total, nsplits=4, block=65536):
    """Start offsets of Hadoop 0.20.2 FileInputFormat.getSplits for one file
    (goal size total/nsplits, SPLIT_SLOP 1.1).  BatchTestFilesTest runs 2 NCs
    x 2 partitions, i.e. getSplits(job, 4) (JobGen.java:67-72); the test HDFS
    block size is 65536 (src/test/resources/hadoop/conf/hdfs-site.xml)."""
    goal = total // nsplits
    size = max(1, min(goal, block))
    starts, rem = [], total
    while rem / size > 1.1:
        starts.append(total - rem)
        rem -= size
    if rem:
        starts.append(total - rem)
    return starts


def test_fixture_count():
    # 8 data/sequence cases + smalltest
    assert len(FIXTURES) == 9


@pytest.mark.parametrize("expected", FIXTURES, ids=[os.path.basename(f) for f in FIXTURES])
def test_fixture_set_level(expected):
    text = open(expected.replace(".expected.txt", ".test.txt"), "rb").read()
    parts = oracle.build_graph(text, _k_of(expected))
    got = oracle.dump_text(parts[0]).splitlines()
    want = open(expected).read().splitlines()
    assert sorted(map(_unordered, got)) == sorted(map(_unordered, want))


def _replay_all(model):
    """Count fixture nodes with a multi-element edge list whose four lists are
    reproduced exactly (order included) by: local aggregation per Hadoop split,
    then cross aggregation over the producers in some order."""
    oracle.set_hash_model(model)
    try:
        total = ok = 0
        for exp in FIXTURES:
            text = open(exp.replace(".expected.txt", ".test.txt"), "rb").read()
            k = _k_of(exp)
            local = []
            for recs in oracle.build_graph(text, k, split_starts=hadoop_splits(len(text)), local_only=True):
                local.append(dict(_lists(l) for l in oracle.dump_text(recs).splitlines()))
            for line in open(exp).read().splitlines():
                key, want = _lists(line)
                if max(len(x) for x in want) < 2:
                    continue
                total += 1
                prods = [d[key] for d in local if key in d]
                for perm in itertools.permutations(prods):
                    cur = [[], [], [], []]
                    for pl in perm:
                        cur = [oracle.union_update(cur[e], pl[e]) for e in range(4)]
                    if cur == want:
                        ok += 1
                        break
        return ok, total
    finally:
        oracle.set_hash_model(0)


def test_fixture_edge_order_replay():
    ok, total = _replay_all(0)
    assert total == 46
    assert ok == total


@pytest.mark.parametrize("model", [1, 2], ids=["jdk8", "jdk6"])
def test_edge_order_replay_rejects_other_hashmaps(model):
    ok, total = _replay_all(model)
    assert ok < total


def test_readids_parse_errors():
    # wrong field count / bad read id / read not longer than k abort the job
    for bad in (b"1\n", b"1\tACGT\tACGT\tACGT\n", b"x\tACGTACGT\n", b"1\tACG\n", b"\n"):
        with pytest.raises(oracle.OracleError):
            oracle.build_graph(bad, 3)
    # readIds must fit 29 bits (ReadHeadInfo.makeUUID's check)
    with pytest.raises(oracle.OracleError):
        oracle.build_graph(b"%d\tACGTACGT\n" % (1 << 29), 3)
    assert oracle.build_graph(b"%d\tACGTACGT\n" % ((1 << 29) - 1), 3)[0]
    # invalid mates are skipped, but still recorded as the other mate's mate
    t = oracle.dump_text(oracle.build_graph(b"7\tACGTTN\tCCGTAA\n", 3)[0])
    assert "7-0_0" not in t and "7-0_1-0 readSeq: CCGTAA mateReadSeq: ACGTTA" in t


def test_partitions_cover_everything():
    text = open(FIXTURES[0].replace(".expected.txt", ".test.txt"), "rb").read()
    for exp in FIXTURES:
        text = open(exp.replace(".expected.txt", ".test.txt"), "rb").read()
        one = sorted(oracle.dump_text(oracle.build_graph(text, 3)[0]).splitlines())
        for P in (2, 3, 4):
            parts = oracle.build_graph(text, 3, nparts=P, split_starts=[0])
            allp = sorted(l for p in parts for l in oracle.dump_text(p).splitlines())
            assert allp == one


@pytest.mark.parametrize("P,crlf", [(1, False), (3, False), (4, True), (8, False)])
def test_threaded_build_equals_split_build(P, crlf):
    """The threaded P-partition build (the multi-core CPU baseline) gives the
    parts of the sequential build with the same producer splits."""
    import random as _r
    rng = _r.Random(P)
    lines = []
    for i in range(400):
        s = "".join(rng.choice("ACGT") for _ in range(rng.randint(25, 60)))
        m = "".join(rng.choice("ACGT") for _ in range(rng.randint(25, 60)))
        lines.append(f"{4 * i + 2}\t{s}\t{m}" if i % 3 else f"{4 * i + 2}\t{s}")
    text = (("\r\n" if crlf else "\n").join(lines) + "\n").encode()
    want = oracle.build_graph(text, 21, P, oracle.split_starts(text, P))
    assert oracle.build_graph_threaded(text, 21, P, 4) == want
