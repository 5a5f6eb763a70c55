"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates the Java reference's graph build on the CPU
(genomix_oracle.cpp).  It is imported only by tests/, by
``__graft_entry__.smoke()`` as the checker, and by bench.py's
``cpu_baseline`` leg.  The product library (genomix_amd) never loads it.

What pins each part of the oracle (no JVM exists here, so the reference
cannot be run; nothing was denied, it is simply unbuildable):

* 2-bit packing, rolling shift, reverse complement, toString:
  KmerFixedTest.java:31-100 vectors (tests/test_oracle_primitives.py) plus
  brute force against string reverse complement for k = 3..63.
* Canonical choice, coverage, read heads and their order, edge SETS:
  the nine current-format text fixtures of genomix-hyracks
  (tests/golden/hyracks/**, k = 3), compared line by line after sorting,
  edge lists as sets (as BatchTestFilesTest does) and read-head sets exactly.
* Edge-list ORDER (JDK-7 HashSet replay): every multi-element edge list in
  those fixtures is reproduced exactly by replaying the fixture input through
  the reference's split/partition structure (four Hadoop input splits, local
  aggregation per split, cross aggregation in some producer order); see
  tests/test_oracle_fixtures.py::test_fixture_edge_order_replay.  Spilled
  external sorts are outside the model (the oracle never spills).
* Record order within a part and partition assignment: closed forms of
  KmerPointable.java:94-107 and KmerPartitionComputerFactory.java:28-53,
  checked against hand-computed Java int arithmetic; parity unpinned by any
  reference test (the reference's tests sort lines).
* Node / ReadHeadSet byte layout: source only (no current-format binary
  fixture exists); round-trip tested.
* SequenceFile container and .crc sidecar: pinned by the reference's binary
  fixtures (tests/golden/seqfile/**): header fields, sync cadence and CRC
  chunks are recomputed from their bytes.
* FASTQ -> readids (readId = 4i+2): source only (GenomixDriver.java:665-714).
* Unitig walker (``unitigs``), pinned three ways
  (tests/test_oracle_merge_pins.py, tests/test_oracle_unitigs.py):
  - a second restatement, ``unitigs_p4``, replays P4's pairwise merges with
    the reference's own primitives, each checked against the reference's
    unit-test vectors: EDGETYPE.resolveEdgeThroughPath
    (NodeFixedTest.java:128-145, all 16 type pairs), Node.mergeEdges
    (NodeRandomTest.java:274-340) and VKmer.mergeWith{FF,FR,RF,RR}Kmer
    (VKmerFixedTest.java:131-452); its bytes equal the walker's on random
    read graphs and complete overlap graphs (k = 3..21, cycles, tandems,
    palindromes);
  - brute force from sequences alone: in graphs whose edges are every
    (k-1)-overlap, each unitig's four edge lists equal the overlaps of the
    oriented unitig sequences; k-mer membership equals a union-find of
    mergeable links;
  - the RandMergeOffsetTest invariant (one genome -> one unitig equal to it
    or its reverse complement, read-head offsets at read positions).
  Planted mistakes (FR/RF swapped in the outer edges, a missing mirror in
  resolveEdgeThroughPath, a missing read-offset flip) each fail one check.
* Graph cleaning (REMOVE_BAD_COVERAGE, TIP_SINGLE_NODE): identity bounds,
  bound invariants and closed forms (tests/test_oracle_clean.py); parity
  unpinned against the reference's one cleaning fixture (retired text
  format, thresholds not in the tree).
* GraphStatistics counters and the FASTA dump: hand-derived counters of a
  single read's k-mer path and of a genome merged into one unitig
  (tests/test_oracle_stats.py).
* BUBBLE: closed forms (tests/test_oracle_bubble.py): a heterozygous SNP or
  indel collapses into the majority allele, giving exactly the merged graph
  of the reads with the minority allele corrected (sequence, coverage, edges
  and read-head offsets; both stored orientations of the two alleles), the
  dissimilarity and length thresholds at their boundaries.  Parity unpinned
  against a reference run: the bubble fixtures (genomix-hadoop
  graphs/bubblemerge) are inputs in the retired text format without outputs.
* FittingMixture: scipy.stats densities for the initial E step, convergence
  to generating parameters, stationarity of the mixture likelihood at the
  fixed point, the cutoff rule, and the closed form of single-valued data
  (tests/test_oracle_mixture.py).
"""

import ctypes
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of the same source (the sanitizer build
# of `make asan`); by default the plain one.
_LIB = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        vp, sz, i32, i64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_int64
        L.ora_last_error.restype = ctypes.c_char_p
        L.ora_build.argtypes = [ctypes.c_char_p, sz, i32, i32, ctypes.POINTER(i64), i32, i32,
                                ctypes.POINTER(vp)]
        L.ora_build_threaded.argtypes = [ctypes.c_char_p, sz, i32, i32, i32, ctypes.POINTER(vp)]
        L.ora_result_nparts.argtypes = [vp]
        L.ora_result_nrecords.argtypes = [vp, i32]
        L.ora_result_nrecords.restype = i64
        L.ora_result_part.argtypes = [vp, i32, ctypes.POINTER(vp)]
        L.ora_result_part.restype = sz
        L.ora_result_free.argtypes = [vp]
        L.ora_free.argtypes = [vp]
        L.ora_seqfile.argtypes = [vp, sz, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                  ctypes.POINTER(vp), ctypes.POINTER(sz)]
        L.ora_crc_sidecar.argtypes = [vp, sz, ctypes.POINTER(vp), ctypes.POINTER(sz)]
        L.ora_dump_text.argtypes = [vp, sz, ctypes.POINTER(vp), ctypes.POINTER(sz)]
        L.ora_fastq_to_readids.argtypes = [ctypes.c_char_p, sz, ctypes.c_char_p, sz, i32,
                                           ctypes.POINTER(vp), ctypes.POINTER(sz)]
        L.ora_pack.argtypes = [ctypes.c_char_p, i32, i32, ctypes.c_char_p]
        L.ora_shift.argtypes = [ctypes.c_char_p, i32, ctypes.c_char]
        L.ora_kmer_str.argtypes = [ctypes.c_char_p, i32, ctypes.c_char_p]
        L.ora_vkmer_hash.argtypes = [ctypes.c_char_p]
        L.ora_vkmer_hash.restype = ctypes.c_int32
        L.ora_union_update.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]
        L.ora_make_uuid.argtypes = [i32, i32, i64, i32]
        L.ora_make_uuid.restype = i64
        L.ora_float_str.argtypes = [ctypes.c_float]
        L.ora_float_str.restype = ctypes.c_char_p
        L.ora_set_hash_model.argtypes = [i32]
        L.ora_unitig_summary.argtypes = [vp, sz, i32, ctypes.POINTER(ctypes.c_double)]
        if hasattr(L, "ora_unitigs"):
            L.ora_unitigs.argtypes = [vp, sz, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)]
        L.ora_unitigs_p4.argtypes = [vp, sz, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)]
        L.ora_set_unitig_plant.argtypes = [i32]
        L.ora_resolve_edge.argtypes = [i32, i32]
        L.ora_vkmer_merge.argtypes = [i32, i32, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]
        L.ora_merge_edges.argtypes = [i32, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, sz]
        for name in ("ora_graph_stats", "ora_dump_fasta"):
            if hasattr(L, name):
                getattr(L, name).argtypes = [vp, sz, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)]
        if hasattr(L, "ora_clean_coverage_then_tips"):
            L.ora_clean_coverage_then_tips.argtypes = [vp, sz, i32, ctypes.c_float, ctypes.c_float, i32,
                                                       ctypes.POINTER(vp), ctypes.POINTER(sz)]
        for name in ("ora_bubble_merge", "ora_bubble_nodes"):
            if hasattr(L, name):
                getattr(L, name).argtypes = [vp, sz, i32, ctypes.c_float, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)]
        if hasattr(L, "ora_clean_pipeline"):
            L.ora_clean_pipeline.argtypes = [vp, sz, i32, ctypes.c_float, ctypes.c_float, i32, ctypes.c_float, i32,
                                             ctypes.POINTER(vp), ctypes.POINTER(sz)]
        if hasattr(L, "ora_fitting_mixture"):
            L.ora_fitting_mixture.argtypes = [vp, sz, ctypes.c_double, i32, ctypes.POINTER(ctypes.c_double)]
        if hasattr(L, "ora_clean_tips"):
            L.ora_clean_tips.argtypes = [vp, sz, i32, i32, ctypes.POINTER(vp), ctypes.POINTER(sz)]
        if hasattr(L, "ora_clean_bad_coverage"):
            L.ora_clean_bad_coverage.argtypes = [vp, sz, i32, ctypes.c_float, ctypes.c_float, ctypes.POINTER(vp),
                                                 ctypes.POINTER(sz)]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise OracleError(rc, lib().ora_last_error().decode(errors="replace"))


def _take(p, n):
    out = ctypes.string_at(p, n.value if hasattr(n, "value") else n)
    lib().ora_free(p)
    return out


def build_graph(text: bytes, k: int, nparts: int = 1, split_starts=None, local_only=False):
    """Run the restated Hyracks graph build; returns a list of per-part record
    streams (i32 recLen, i32 keyLen, VKmer key, Node value; no header)."""
    L = lib()
    res = ctypes.c_void_p()
    if split_starts is None:
        sp, ns = None, 0
    else:
        arr = (ctypes.c_int64 * len(split_starts))(*split_starts)
        sp, ns = arr, len(split_starts)
    _check(L.ora_build(text, len(text), k, nparts, sp, ns, 1 if local_only else 0, ctypes.byref(res)))
    try:
        parts = []
        for p in range(L.ora_result_nparts(res)):
            d = ctypes.c_void_p()
            n = L.ora_result_part(res, p, ctypes.byref(d))
            parts.append(ctypes.string_at(d, n))
        return parts
    finally:
        L.ora_result_free(res)


def split_starts(text: bytes, P: int):
    """Line-aligned starts of P equal byte ranges (the producers of
    build_graph_threaded)."""
    n = len(text)
    st = [0]
    for p in range(1, P):
        s = max(st[-1], n * p // P)
        # a split starts right after a line terminator ("\r\n" counts as one)
        while s > st[-1] and s < n and (text[s - 1] not in (10, 13) or (text[s - 1] == 13 and text[s] == 10)):
            s += 1
        st.append(s)
    return st


def build_graph_threaded(text: bytes, k: int, nparts: int, nthreads: int):
    """The P-partition build with its producer and consumer stages on
    threads (Hyracks with threadsPerMachine = P); equals
    build_graph(text, k, P, split_starts(text, P))."""
    L = lib()
    res = ctypes.c_void_p()
    _check(L.ora_build_threaded(text, len(text), k, nparts, nthreads, ctypes.byref(res)))
    try:
        parts = []
        for p in range(L.ora_result_nparts(res)):
            d = ctypes.c_void_p()
            n = L.ora_result_part(res, p, ctypes.byref(d))
            parts.append(ctypes.string_at(d, n))
        return parts
    finally:
        L.ora_result_free(res)


KEY_CLASS = "edu.uci.ics.genomix.data.types.VKmer"
VALUE_CLASS = "edu.uci.ics.genomix.data.types.Node"


def seqfile(records: bytes, sync: bytes, key_class=KEY_CLASS, value_class=VALUE_CLASS) -> bytes:
    """Wrap a record stream in a SequenceFile v6 container (header + syncs)."""
    assert len(sync) == 16
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_seqfile(records, len(records), sync, key_class.encode(), value_class.encode(),
                             ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def crc_sidecar(data: bytes) -> bytes:
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_crc_sidecar(data, len(data), ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def dump_text(records: bytes) -> str:
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_dump_text(records, len(records), ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n).decode()


def fastq_to_readids(mate1: bytes, mate2: bytes = None) -> bytes:
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    paired = mate2 is not None
    _check(lib().ora_fastq_to_readids(mate1, len(mate1), mate2 or b"", len(mate2 or b""), int(paired),
                                      ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def pack(s: str, k: int, reverse=False) -> bytes:
    nb = (k + 3) // 4
    out = ctypes.create_string_buffer(nb)
    _check(lib().ora_pack(s.encode(), k, int(reverse), out))
    return out.raw[:nb]


def shift(b: bytes, k: int, c: str) -> bytes:
    buf = ctypes.create_string_buffer(bytes(b), len(b))
    _check(lib().ora_shift(buf, k, c.encode()))
    return buf.raw[:len(b)]


def kmer_str(b: bytes, k: int) -> str:
    out = ctypes.create_string_buffer(k + 1)
    _check(lib().ora_kmer_str(bytes(b), k, out))
    return out.value.decode()


def vkmer_hash(s: str) -> int:
    return lib().ora_vkmer_hash(s.encode())


def union_update(L, M):
    out = ctypes.create_string_buffer(4096)
    _check(lib().ora_union_update(",".join(L).encode(), ",".join(M).encode(), out, 4096))
    v = out.value.decode()
    return v.split(",") if v else []


def set_hash_model(model: int):
    """Test hook: 0 = JDK-7 HashSet (the model), 1 = JDK-8, 2 = JDK-6 (planted
    wrong models, negative controls only)."""
    lib().ora_set_hash_model(model)


def make_uuid(mate, lib_id, read_id, offset):
    v = lib().ora_make_uuid(mate, lib_id, read_id, offset)
    if v == -(1 << 63):
        raise OracleError(-5, lib().ora_last_error().decode())
    return v


def float_str(f: float) -> str:
    return lib().ora_float_str(f).decode()


def unitigs(records: bytes, k: int) -> bytes:
    """Sequential maximal-unitig walker over a graph record stream; returns one
    merged record per maximal unitig (key = its smallest canonical k-mer,
    sequence oriented so that k-mer reads forward), sorted by key."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_unitigs(records, len(records), k, ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def unitigs_p4(records: bytes, k: int) -> bytes:
    """The same unitigs as ``unitigs``, computed by replaying P4's pairwise
    merges (Node.mergeWithNode, the neighbour updates and
    resolveEdgeThroughPath) one node at a time into each path's survivor."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_unitigs_p4(records, len(records), k, ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def set_unitig_plant(mode: int):
    """Test hook: 0 = the restatement; 1 = FR/RF swapped in the walker's
    outer edges; 2 = resolveEdgeThroughPath without its mirror; 3 = the fold's
    read-id merge without the offset flip (negative controls only)."""
    lib().ora_set_unitig_plant(mode)


ETYPES = ("FF", "FR", "RF", "RR")


def resolve_edge(b_to_a: str, b_to_c: str):
    """EDGETYPE.resolveEdgeThroughPath; None for an invalid path."""
    r = lib().ora_resolve_edge(ETYPES.index(b_to_a), ETYPES.index(b_to_c))
    return None if r < 0 else ETYPES[r]


def vkmer_merge(edge_type: str, initial: int, this: str, other: str) -> str:
    """VKmer.mergeWith<edge_type>Kmer(initial, other) applied to ``this``."""
    out = ctypes.create_string_buffer(len(this) + len(other) + 2)
    _check(lib().ora_vkmer_merge(ETYPES.index(edge_type), initial, this.encode(), other.encode(), out, len(out)))
    return out.value.decode()


def merge_edges(edge_type: str, this: dict, other: dict) -> dict:
    """Node.mergeEdges(edge_type, other) on {type: [kmer strings]} lists;
    returns this node's lists afterwards."""
    enc = lambda d: "|".join(",".join(d.get(t, [])) for t in ETYPES)  # noqa: E731
    out = ctypes.create_string_buffer(4096)
    _check(lib().ora_merge_edges(ETYPES.index(edge_type), enc(this).encode(), enc(other).encode(), out, 4096))
    fields = out.value.decode().split("|")
    return {t: (f.split(",") if f else []) for t, f in zip(ETYPES, fields)}


def clean_bad_coverage(records: bytes, k: int, min_cov=3.0, max_cov=float("inf")) -> bytes:
    """REMOVE_BAD_COVERAGE then MERGE (RemoveBadCoverageVertex.java:38-60;
    defaults GenomixJobConf.java:498-502) on a graph record stream: merged
    paths whose average coverage lies outside [min_cov, max_cov] are
    removed with their nodes, survivors lose the edges to them, and the
    surviving graph is merged again.  Returns the merged records."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_clean_bad_coverage(records, len(records), k, ctypes.c_float(min_cov), ctypes.c_float(max_cov),
                                        ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def clean_tips(records: bytes, k: int, max_len=None) -> bytes:
    """TIP_SINGLE_NODE then MERGE (SingleNodeTipRemoveVertex.java:47-85;
    default max_len = k, GenomixJobConf.java:504-505) on a graph record
    stream: merged vertices with one edge on one side and none on the other
    and at most max_len letters are removed, their neighbours drop the edge,
    and the surviving graph is merged again.  Returns the merged records."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_clean_tips(records, len(records), k, k if max_len is None else max_len, ctypes.byref(p),
                                ctypes.byref(n)))
    return _take(p, n)


def clean_coverage_then_tips(records: bytes, k: int, min_cov=3.0, max_cov=float("inf"), max_len=None) -> bytes:
    """The reference pipeline's cleaning prefix (GenomixJobConf.java:511-517):
    REMOVE_BAD_COVERAGE, MERGE, TIP_SINGLE_NODE, MERGE."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_clean_coverage_then_tips(records, len(records), k, ctypes.c_float(min_cov),
                                              ctypes.c_float(max_cov), k if max_len is None else max_len,
                                              ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def graph_stats(records: bytes, k: int) -> dict:
    """GraphStatistics counters (GraphStatistics.java:79-146) of a record
    stream: {"totals": {...}, "maximum": {...}, "<name>-bins": {value: n}}."""
    import json
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_graph_stats(records, len(records), k, ctypes.byref(p), ctypes.byref(n)))
    return json.loads(_take(p, n).decode())


def dump_fasta(records: bytes, k: int) -> bytes:
    """DriverUtils.dumpGraph (DriverUtils.java:109-114) of a record stream."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_dump_fasta(records, len(records), k, ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def bubble_merge(records: bytes, k: int, max_dissimilarity=0.05, max_length=None) -> bytes:
    """BUBBLE then MERGE (SimpleBubbleMergeVertex.java:51-320; defaults
    GenomixJobConf.java:478-482: 0.05 and 5k) on a node-graph record stream;
    returns the merged records."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_bubble_merge(records, len(records), k, ctypes.c_float(max_dissimilarity),
                                  -1 if max_length is None else max_length, ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def clean_pipeline(records: bytes, k: int, min_cov=3.0, max_cov=float("inf"), tip_length=None,
                   max_dissimilarity=0.05, max_length=None) -> bytes:
    """The reference pipeline's cleaning after BUILD (GenomixJobConf.java:511-517):
    REMOVE_BAD_COVERAGE, MERGE, TIP_SINGLE_NODE, MERGE, BUBBLE, MERGE."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_clean_pipeline(records, len(records), k, ctypes.c_float(min_cov), ctypes.c_float(max_cov),
                                    -1 if tip_length is None else tip_length, ctypes.c_float(max_dissimilarity),
                                    -1 if max_length is None else max_length, ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def bubble_nodes(records: bytes, k: int, max_dissimilarity=0.05, max_length=None) -> bytes:
    """The node graph after BUBBLE (before the merge that follows)."""
    p, n = ctypes.c_void_p(), ctypes.c_size_t()
    _check(lib().ora_bubble_nodes(records, len(records), k, ctypes.c_float(max_dissimilarity),
                                  -1 if max_length is None else max_length, ctypes.byref(p), ctypes.byref(n)))
    return _take(p, n)


def fitting_mixture(data, max_value, iterations=10) -> dict:
    """FittingMixture.fittingMixture (FittingMixture.java:89-219) over the
    expanded coverage data (one value per node).  Returns the cutoff and the
    fitted exponential mean, normal mean / sd and priors."""
    import numpy as np
    d = np.ascontiguousarray(data, dtype=np.float64)
    a = (ctypes.c_double * 6)()
    _check(lib().ora_fitting_mixture(d.ctypes.data, d.size, float(max_value), iterations, a))
    return dict(zip(MIXTURE_KEYS, list(a)))


MIXTURE_KEYS = ["cutoff", "exp_mean", "normal_mean", "normal_std", "prob_exp", "prob_normal"]


def unitig_summary(records: bytes, k: int):
    """(records, total k-mers, sum of coverage x length, keys ascending) of a
    merged record stream, decoded by the oracle's own Node reader."""
    a = (ctypes.c_double * 4)()
    _check(lib().ora_unitig_summary(records, len(records), k, a))
    return int(a[0]), int(a[1]), a[2], bool(a[3])
