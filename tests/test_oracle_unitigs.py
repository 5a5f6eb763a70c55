"""The oracle's unitig walker against the invariants the reference tests.

RandMergeOffsetTest (genomix-driver/.../pipelinetests/RandMergeOffsetTest.java:94-174)
builds and merges one generated read: the merged node equals the source or
its reverse complement and every read head sits at its read's position.
Beyond that, unitig membership is checked against a brute-force
construction on small random graphs.
"""

import random
import struct

import pytest

from oracle import oracle

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def parse_records(rec):
    """Minimal Node decoder (Node.java:408-487) for the checks below."""
    out = []
    off = 0
    while off < len(rec):
        n, kl = struct.unpack(">ii", rec[off:off + 8])
        body = rec[off + 8:off + 8 + n]
        out.append((body[:kl], body[kl:]))
        off += 8 + n
    return out


def lines_of(records):
    return oracle.dump_text(records).splitlines()


def field(line, name):
    """One field of a text-dumped Node: KEY\\t{FF:..\\tFR:..\\tRF:..\\tRR:..\\t5':[..], ~5':[..]\\tkmer:..\\tcov:..}"""
    parts = line.split("\t", 1)[1][1:-1].split("\t")
    heads = parts[4]
    a, b = heads.split(", ~5':")
    vals = {"FF": parts[0][3:], "FR": parts[1][3:], "RF": parts[2][3:], "RR": parts[3][3:],
            "5'": a[len("5':"):], "~5'": b, "kmer": parts[5][5:], "cov": parts[6][4:]}
    return vals[name]


def tiled_reads(genome, L, step, rng):
    reads = []
    for p in range(0, len(genome) - L + 1, step):
        s = genome[p:p + L]
        fwd = rng.random() < 0.5
        reads.append((p, fwd, s if fwd else rc(s)))
    last = len(genome) - L
    if reads[-1][0] != last:
        s = genome[last:]
        reads.append((last, True, s))
    return reads


def _genome_without_repeats(rng, G, k):
    """A random genome whose (k-1)-mers and their reverse complements are all
    distinct (bounded number of draws)."""
    for _ in range(200):
        genome = "".join(rng.choice("ACGT") for _ in range(G))
        ks = [genome[i:i + k - 1] for i in range(G - k + 2)]
        allk = ks + [rc(x) for x in ks]
        if len(set(allk)) == len(allk):
            return genome
    raise AssertionError("no repeat-free genome found")


@pytest.mark.parametrize("k,G", [(9, 300), (21, 1200), (31, 1200), (33, 1200), (55, 1200)])
def test_single_genome_merges_to_one_unitig(k, G):
    rng = random.Random(k)
    genome = _genome_without_repeats(rng, G, k)
    reads = tiled_reads(genome, 100, 7, rng)
    text = "".join(f"{4 * i + 2}\t{s}\n" for i, (_, _, s) in enumerate(reads)).encode()
    graph = oracle.build_graph(text, k)[0]
    merged = oracle.unitigs(graph, k)
    ls = lines_of(merged)
    assert len(ls) == 1
    seq = field(ls[0], "kmer")
    assert seq in (genome, rc(genome))
    forward = seq == genome
    # read heads: readId-offset_mate-lib; offset = first k-mer position in the
    # unitig, + k-1 when the read runs against the unitig
    heads = {}
    for name, flipped in (("5'", False), ("~5'", True)):
        v = field(ls[0], name)
        if v == "null":
            continue
        for h in v[1:-1].split(","):
            rid, rest = h.split("-", 1)
            off = int(rest.split("_")[0])
            heads[int(rid)] = (off, flipped)
    assert len(heads) == len(reads)
    for i, (p, fwd, s) in enumerate(reads):
        off, flipped = heads[4 * i + 2]
        same = fwd == forward
        q = p if forward else G - (p + 100)  # read start in unitig coordinates
        if same:
            assert (off, flipped) == (q, False)
        else:
            assert (off, flipped) == (q + 100 - k + k - 1, True)
    # coverage = mean k-mer count
    cov = float(field(ls[0], "cov")[:-1])
    nk = G - k + 1
    assert abs(cov - sum(100 - k + 1 for _ in reads) / nk) < 1e-3 * cov


def _brute_unitigs(records, k):
    """Components of mergeable links computed by plain set logic."""
    text = lines_of(records)
    nodes = {}
    for line in text:
        key = line.split("\t")[0]
        lists = {}
        for name in ("FF", "FR", "RF", "RR"):
            v = field(line, name)
            lists[name] = [] if v == "null" else v[1:-1].split(",")
        nodes[key] = lists
    side = {"FF": "F", "FR": "F", "RF": "R", "RR": "R"}
    mirror = {"FF": "RR", "RR": "FF", "FR": "FR", "RF": "RF"}

    def deg(x, s):
        return sum(len(v) for t, v in nodes[x].items() if side[t] == s)

    def tandem(x):
        return any(x in v for v in nodes[x].values())

    parent = {x: x for x in nodes}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for x, lists in nodes.items():
        for t, v in lists.items():
            if len(v) == 1 and deg(x, side[t]) == 1:
                y = v[0]
                if y != x and not tandem(x) and not tandem(y) and deg(y, side[mirror[t]]) == 1:
                    parent[find(x)] = find(y)
    comps = {}
    for x in nodes:
        comps.setdefault(find(x), set()).add(x)
    return sorted(sorted(c) for c in comps.values())


@pytest.mark.parametrize("seed", range(6))
def test_unitig_membership_brute_force(seed):
    rng = random.Random(100 + seed)
    k = rng.choice([3, 4, 5, 7])
    G = rng.randint(30, 120)
    genome = "".join(rng.choice("ACGT") for _ in range(G))
    reads = []
    for i in range(rng.randint(5, 30)):
        L = rng.randint(k + 1, 25)
        p = rng.randint(0, max(0, G - L))
        s = genome[p:p + L]
        if rng.random() < 0.5:
            s = rc(s)
        if rng.random() < 0.3:
            j = rng.randrange(len(s))
            s = s[:j] + rng.choice("ACGT") + s[j + 1:]
        reads.append(s)
    text = "".join(f"{4 * i + 2}\t{s}\n" for i, s in enumerate(reads)).encode()
    graph = oracle.build_graph(text, k)[0]
    merged = oracle.unitigs(graph, k)
    want = _brute_unitigs(graph, k)
    # membership: every unitig's k-mers (canonical) form one brute-force component
    canon_of = {}
    for line in lines_of(graph):
        key = line.split("\t")[0]
        canon_of[key] = key
        canon_of[rc(key)] = key
    got = []
    for line in lines_of(merged):
        seq = field(line, "kmer")
        key = line.split("\t")[0]
        if seq == "null":
            got.append([key])
        else:
            got.append(sorted({canon_of[seq[i:i + k]] for i in range(len(seq) - k + 1)}))
    assert sorted(got) == want
