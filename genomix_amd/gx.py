"""ctypes binding of libgenomix_gx.so (include/genomix_gx.h).

``build_graph`` plays the part of GenomixHyracksDriver.runJob(conf,
Plan.BUILD_DEBRUIJN_GRAPH) (GenomixHyracksDriver.java:84-135) for one task:
readids text in, ``part-<p>`` record streams out.  There is no CPU path: if
the HIP library or a HIP device is missing, calls fail with GxError.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgenomix_gx.so")

GX_OK = 0
GX_EINVAL = -22
GX_EINPUT = -5
GX_ENOMEM = -12
GX_EDEVICE = -19
GX_ESTATE = -16
GX_ENOTSUP = -95
GX_FMT_READIDS = 0
GX_FMT_FASTQ = 1
EDGE_ORDER = {"canonical": 0, "jdk7": 1}
METHOD = {"sort": 0, "hash": 1}  # GX_AGG_SORT, GX_AGG_HASH

EXPORTED = [
    "gx_build_create", "gx_build_add_reads", "gx_build_add_reads_device", "gx_build_finish",
    "gx_build_part_info", "gx_build_emit_part", "gx_build_part_device", "gx_build_stats",
    "gx_build_reset", "gx_last_error", "gx_build_destroy", "gx_build_set_profiling",
    "gx_build_phase_times", "gx_build_phase_name", "gx_unitig_run", "gx_unitig_info", "gx_unitig_emit",
    "gx_build_add_fastq_pair", "gx_build_readids",
    "gx_mlp_create", "gx_mlp_num_params", "gx_mlp_set_params", "gx_mlp_get_params", "gx_mlp_forward",
    "gx_mlp_gradient", "gx_mlp_train_step", "gx_mlp_last_error", "gx_mlp_destroy", "gx_build_set_sizer",
    "gx_build_features", "gx_build_set_tuning", "gx_build_set_config", "gx_batch_create", "gx_batch_set",
    "gx_batch_set_sizer", "gx_batch_set_tuning", "gx_batch_run", "gx_batch_set_stats", "gx_batch_last_error",
    "gx_batch_destroy", "gx_build_set_method", "gx_build_set_stream", "gx_clean_bad_coverage", "gx_clean_tips",
    "gx_graph_stats", "gx_graph_stat_totals", "gx_graph_stat", "gx_graph_stat_bins", "gx_dump_fasta",
    "gx_coverage_cutoff", "gx_fit_mixture", "gx_clean_bubbles", "gx_batch_retried", "gx_unitig_limits", "gx_unitig_set_ranking",
    "gx_mlp_gradient_noalloc", "gx_batch_set_methods",
]
GRAPHS = {"nodes": 0, "unitigs": 1}
STAT_NAMES = ["degree", "kmerLength", "coverage", "unflippedReadIds", "flippedReadIds", "scaffoldSeedScore",
              "kmerLength-with-FORWARD", "coverage-with-FORWARD", "scaffoldSeedScore-with-FORWARD",
              "kmerLength-with-REVERSE", "coverage-with-REVERSE", "scaffoldSeedScore-with-REVERSE"]
MIXTURE_KEYS = ["cutoff", "exp_mean", "normal_mean", "normal_std", "prob_exp", "prob_normal"]
TOTAL_NAMES = ["nodes", "pathNode", "tips-FORWARD", "tips-REVERSE", "tips-BOTH", "tips-ONE", "selfEdge-FF",
               "selfEdge-FR", "selfEdge-RF", "selfEdge-RR"]


class GxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"genomix_gx error {code}: {msg}")
        self.code = code


class _Cfg(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int), ("nparts", ctypes.c_int), ("max_readids_per_edge", ctypes.c_int),
                ("edge_order", ctypes.c_int), ("device", ctypes.c_int), ("table_hint", ctypes.c_size_t)]


_lib = None
BATCH_CONSUMER = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)
PART_SINK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                             ctypes.c_uint64)


def lib():
    """Load the in-tree HIP library (never a fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GxError(GX_EDEVICE, f"{LIB_PATH} is not built; run genomix_amd/build.py (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i32, u64p = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)
        L.gx_build_create.argtypes = [ctypes.POINTER(_Cfg), ctypes.POINTER(vp)]
        L.gx_build_add_reads.argtypes = [vp, ctypes.c_char_p, sz, i32]
        L.gx_build_add_reads_device.argtypes = [vp, vp, sz, i32]
        L.gx_build_finish.argtypes = [vp]
        L.gx_build_part_info.argtypes = [vp, i32, u64p, u64p]
        L.gx_build_emit_part.argtypes = [vp, i32, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        L.gx_build_part_device.argtypes = [vp, i32, ctypes.POINTER(vp), u64p]
        L.gx_build_stats.argtypes = [vp, u64p, i32]
        L.gx_build_reset.argtypes = [vp]
        L.gx_last_error.argtypes = [vp]
        L.gx_last_error.restype = ctypes.c_char_p
        L.gx_build_destroy.argtypes = [vp]
        L.gx_build_set_profiling.argtypes = [vp, i32]
        L.gx_build_phase_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double), i32]
        L.gx_build_phase_name.argtypes = [i32]
        L.gx_build_phase_name.restype = ctypes.c_char_p
        L.gx_unitig_run.argtypes = [vp]
        L.gx_unitig_info.argtypes = [vp, u64p, u64p]
        L.gx_unitig_limits.argtypes = [u64p, u64p]
        L.gx_unitig_set_ranking.argtypes = [vp, i32]
        L.gx_unitig_emit.argtypes = [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        L.gx_build_add_fastq_pair.argtypes = [vp, ctypes.c_char_p, sz, ctypes.c_char_p, sz]
        L.gx_build_readids.argtypes = [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        fp = ctypes.POINTER(ctypes.c_float)
        L.gx_mlp_create.argtypes = [i32, i32, i32, i32, ctypes.POINTER(vp)]
        L.gx_mlp_num_params.argtypes = [vp]
        L.gx_mlp_num_params.restype = sz
        L.gx_mlp_set_params.argtypes = [vp, fp, sz]
        L.gx_mlp_get_params.argtypes = [vp, fp, sz]
        L.gx_mlp_forward.argtypes = [vp, fp, i32, fp]
        L.gx_mlp_gradient.argtypes = [vp, fp, fp, i32, fp, fp]
        L.gx_mlp_gradient_noalloc.argtypes = [vp, vp, vp, i32, vp, vp, vp]
        L.gx_mlp_train_step.argtypes = [vp, fp, fp, i32, ctypes.c_float, fp]
        L.gx_mlp_last_error.argtypes = [vp]
        L.gx_mlp_last_error.restype = ctypes.c_char_p
        L.gx_mlp_destroy.argtypes = [vp]
        L.gx_build_set_sizer.argtypes = [vp, vp]
        L.gx_build_features.argtypes = [vp, fp, i32, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_double)]
        L.gx_build_set_tuning.argtypes = [vp, i32, ctypes.c_float]
        L.gx_build_set_config.argtypes = [vp, ctypes.POINTER(_Cfg)]
        L.gx_build_set_method.argtypes = [vp, i32]
        L.gx_build_set_stream.argtypes = [vp, ctypes.c_uint64, PART_SINK, vp]
        L.gx_clean_bad_coverage.argtypes = [vp, ctypes.c_float, ctypes.c_float]
        L.gx_clean_tips.argtypes = [vp, i32]
        L.gx_clean_bubbles.argtypes = [vp, ctypes.c_float, i32]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.gx_graph_stats.argtypes = [vp, i32]
        L.gx_graph_stat_totals.argtypes = [vp, u64p, i32]
        L.gx_graph_stat.argtypes = [vp, i32, i64p, i64p, u64p]
        L.gx_graph_stat_bins.argtypes = [vp, i32, i64p, u64p, sz]
        L.gx_dump_fasta.argtypes = [vp, i32, ctypes.c_char_p, sz, ctypes.POINTER(sz)]
        dp = ctypes.POINTER(ctypes.c_double)
        L.gx_coverage_cutoff.argtypes = [vp, i32, dp]
        L.gx_fit_mixture.argtypes = [dp, u64p, sz, ctypes.c_double, i32, dp]
        L.gx_batch_create.argtypes = [i32, i32, ctypes.POINTER(vp)]
        L.gx_batch_set.argtypes = [vp, i32, ctypes.POINTER(_Cfg), vp, sz, i32, i32]
        L.gx_batch_set_sizer.argtypes = [vp, vp]
        L.gx_batch_set_tuning.argtypes = [vp, i32, i32, u64p, ctypes.POINTER(ctypes.c_int), fp]
        L.gx_batch_set_methods.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_int)]
        L.gx_batch_run.argtypes = [vp, i32, BATCH_CONSUMER, vp]
        L.gx_batch_set_stats.argtypes = [vp, i32, u64p, i32, ctypes.POINTER(ctypes.c_double)]
        L.gx_batch_last_error.argtypes = [vp]
        L.gx_batch_retried.argtypes = [vp]
        L.gx_batch_last_error.restype = ctypes.c_char_p
        L.gx_batch_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _host_bytes(n, fill) -> bytes:
    """n bytes written by fill(char pointer) -> count written, as bytes.  The
    buffer is not zero-filled first, and the result is its one copy."""
    import numpy as np
    arr = np.empty(max(int(n), 1), dtype=np.uint8)
    w = fill(ctypes.cast(ctypes.c_void_p(arr.ctypes.data), ctypes.c_char_p))
    return arr[:w].tobytes()


class GraphBuilder:
    """One graph-build task on one GPU (a gx_build handle)."""

    def __init__(self, k, nparts=1, edge_order="jdk7", device=0, table_hint=0, max_readids_per_edge=250,
                 sizer=None, method=None):
        self._h = ctypes.c_void_p()
        self._sizer = None
        cfg = _Cfg(k, nparts, max_readids_per_edge, EDGE_ORDER[edge_order], device, table_hint)
        rc = lib().gx_build_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != 0:
            raise GxError(rc, lib().gx_last_error(None).decode())
        self.k = k
        self.nparts = nparts
        if sizer is not None:
            self.set_sizer(sizer)
        if method is not None:
            self.set_method(method)

    def set_method(self, method):
        """'sort' (default: sort-partitioned aggregation) or 'hash' (atomic
        hash table, sized by the learned predictor when one is attached)."""
        self._check(lib().gx_build_set_method(self._h, METHOD[method]))

    def set_stream(self, range_occurrences=0, sink=None):
        """Key-range passes (include/genomix_stream.h): at most about
        `range_occurrences` k-mer occurrences per pass (0: as HBM needs) and,
        with `sink`, each range's records handed to
        ``sink(part, device_address, nbytes, nrec)`` instead of kept."""
        self._sink_exc = None
        if sink is None:
            cb = ctypes.cast(None, PART_SINK)
        else:
            def call(_user, part, addr, nbytes, nrec):
                try:
                    return int(sink(part, addr, nbytes, nrec) or 0)
                except BaseException as e:  # reported by finish()
                    self._sink_exc = e
                    return GX_EINVAL
            cb = PART_SINK(call)
        self._check(lib().gx_build_set_stream(self._h, range_occurrences, cb, None))
        self._sink_cb = cb  # alive while the library holds it

    def set_sizer(self, sizer):
        """Attach a trained sizing network (genomix_amd.sizer.Mlp) or None."""
        self._check(lib().gx_build_set_sizer(self._h, sizer.handle if sizer is not None else None))
        self._sizer = sizer  # keep it alive while attached

    def features(self):
        """(features, HLL registers, [predicted distinct, predicted skew]) of the last finish."""
        import numpy as np
        f = np.zeros(72, np.float32)
        r = np.zeros(4096, np.uint8)
        p = (ctypes.c_double * 2)()
        self._check(lib().gx_build_features(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 72,
                                            r.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), p))
        return f, r, [p[0], p[1]]

    def _check(self, rc):
        if rc != 0:
            raise GxError(rc, lib().gx_last_error(self._h).decode(errors="replace"))

    def add_reads(self, text: bytes):
        self._check(lib().gx_build_add_reads(self._h, text, len(text), GX_FMT_READIDS))

    def add_reads_device(self, ptr: int, nbytes: int):
        self._check(lib().gx_build_add_reads_device(self._h, ctypes.c_void_p(ptr), nbytes, GX_FMT_READIDS))

    def add_fastq(self, mate1: bytes, mate2: bytes = None):
        """FASTQ bytes of one library (single-end, or both files of a pair);
        converted to readids lines on the device at finish."""
        if mate2 is None:
            self._check(lib().gx_build_add_reads(self._h, mate1, len(mate1), GX_FMT_FASTQ))
        else:
            self._check(lib().gx_build_add_fastq_pair(self._h, mate1, len(mate1), mate2, len(mate2)))

    def finish(self):
        rc = lib().gx_build_finish(self._h)
        exc = getattr(self, "_sink_exc", None)
        if exc is not None:
            self._sink_exc = None
            raise exc
        self._check(rc)

    def readids(self) -> bytes:
        """The library-N.readids text made from FASTQ input (after finish)."""
        n = ctypes.c_size_t()
        self._check(lib().gx_build_readids(self._h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(max(n.value, 1))
        self._check(lib().gx_build_readids(self._h, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def part_info(self, part=0):
        n, b = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().gx_build_part_info(self._h, part, ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    def part_bytes(self, part=0) -> bytes:
        _, nb = self.part_info(part)

        def fill(buf):
            w = ctypes.c_size_t()
            self._check(lib().gx_build_emit_part(self._h, part, buf, nb, ctypes.byref(w)))
            return w.value
        return _host_bytes(nb, fill)

    def set_ranking(self, mode):
        """"auto" (walks for odd k, pointer jumping otherwise) or
        "pointer_jumping"; the records are the same."""
        self._check(lib().gx_unitig_set_ranking(self._h, {"auto": 0, "pointer_jumping": 1}[mode]))

    def run_unitigs(self):
        """Unitig compression of the built graph; records stay on the device."""
        self._check(lib().gx_unitig_run(self._h))

    def unitig_count(self):
        n, nb = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().gx_unitig_info(self._h, ctypes.byref(n), ctypes.byref(nb)))
        return n.value

    def unitigs(self) -> bytes:
        """Unitig compression of the built graph; returns the merged records."""
        self._check(lib().gx_unitig_run(self._h))
        return self.unitig_records()

    def unitig_records(self) -> bytes:
        """The merged records of the last unitig pass or cleaning step."""
        n, nb = ctypes.c_uint64(), ctypes.c_uint64()
        self._check(lib().gx_unitig_info(self._h, ctypes.byref(n), ctypes.byref(nb)))

        def fill(buf):
            w = ctypes.c_size_t()
            self._check(lib().gx_unitig_emit(self._h, buf, nb.value, ctypes.byref(w)))
            return w.value
        return _host_bytes(nb.value, fill)

    def clean_bad_coverage(self, min_cov=3.0, max_cov=float("inf"), emit=True):
        """REMOVE_BAD_COVERAGE then MERGE (include/genomix_clean.h) on the
        graph after a unitig pass; returns the merged records (copied to the
        host; emit=False leaves them in HBM and returns None)."""
        self._check(lib().gx_clean_bad_coverage(self._h, ctypes.c_float(min_cov), ctypes.c_float(max_cov)))
        return self.unitig_records() if emit else None

    def clean_tips(self, max_len=-1, emit=True):
        """TIP_SINGLE_NODE then MERGE (include/genomix_clean.h); max_len < 0:
        k.  Returns the merged records (emit=False: None)."""
        self._check(lib().gx_clean_tips(self._h, max_len))
        return self.unitig_records() if emit else None

    def clean_bubbles(self, max_dissimilarity=0.05, max_len=-1, emit=True):
        """BUBBLE then MERGE (include/genomix_clean.h); max_len < 0: 5k.
        Returns the merged records (emit=False: None)."""
        self._check(lib().gx_clean_bubbles(self._h, ctypes.c_float(max_dissimilarity), max_len))
        return self.unitig_records() if emit else None

    def graph_stats(self, graph="unitigs") -> dict:
        """GraphStatistics counters (include/genomix_stats.h) of the node graph
        or of the last merge, laid out as the reference's counter groups:
        {"totals": {...}, "maximum": {...}, "<name>-bins": {value: count}}
        (counters that stay 0 are left out, as Hadoop never creates them)."""
        self._check(lib().gx_graph_stats(self._h, GRAPHS[graph]))
        tot = (ctypes.c_uint64 * len(TOTAL_NAMES))()
        self._check(lib().gx_graph_stat_totals(self._h, tot, len(TOTAL_NAMES)))
        out = {"totals": {n: v for n, v in zip(TOTAL_NAMES, tot) if v}, "maximum": {}}
        for sid, name in enumerate(STAT_NAMES):
            total, mx, nb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_uint64()
            self._check(lib().gx_graph_stat(self._h, sid, ctypes.byref(total), ctypes.byref(mx), ctypes.byref(nb)))
            if nb.value == 0:
                continue
            vals = (ctypes.c_int64 * nb.value)()
            cnts = (ctypes.c_uint64 * nb.value)()
            self._check(lib().gx_graph_stat_bins(self._h, sid, vals, cnts, nb.value))
            out[name + "-bins"] = {str(v): c for v, c in zip(vals, cnts)}
            out["totals"][name] = total.value
            out["maximum"][name] = mx.value
        return {g: c for g, c in out.items() if c}  # a group with no counter does not exist

    def coverage_cutoff(self, graph="nodes", iterations=10) -> dict:
        """The driver's coverage cutoff (include/genomix_stats.h,
        GenomixDriver.java:119-136): statistics of `graph` (GenomixDriver uses
        the build's node graph), then the exponential + normal mixture fit of
        their coverage counters.  Returns MIXTURE_KEYS -> value; a positive
        "cutoff" is the min_cov GenomixDriver hands to REMOVE_BAD_COVERAGE."""
        self._check(lib().gx_graph_stats(self._h, GRAPHS[graph]))
        a = (ctypes.c_double * len(MIXTURE_KEYS))()
        self._check(lib().gx_coverage_cutoff(self._h, iterations, a))
        return dict(zip(MIXTURE_KEYS, list(a)))

    def fasta(self, graph="unitigs") -> bytes:
        """DriverUtils.dumpGraph of the node graph or of the last merge."""
        n = ctypes.c_size_t()
        self._check(lib().gx_dump_fasta(self._h, GRAPHS[graph], None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(max(n.value, 1))
        self._check(lib().gx_dump_fasta(self._h, GRAPHS[graph], buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def stats(self):
        a = (ctypes.c_uint64 * 16)()
        self._check(lib().gx_build_stats(self._h, a, 16))
        keys = ["occurrences", "distinct", "heads", "table_capacity", "replayed_nodes", "replayed_occurrences",
                "max_coverage", "predicted_distinct", "table_rebuilds", "oversized_bucket_records", "pj_rounds",
                "key_ranges", "pj_successor_reads", "walkers", "device_allocations", "buffers_reclaimed"]
        return dict(zip(keys, list(a)))

    def set_profiling(self, on=True):
        self._check(lib().gx_build_set_profiling(self._h, int(on)))

    def phase_times(self):
        """Device milliseconds per phase of the last finish (hipEvents)."""
        a = (ctypes.c_double * 16)()
        n = lib().gx_build_phase_times(self._h, a, 16)
        return {lib().gx_build_phase_name(i).decode(): a[i] for i in range(n)}

    def reset(self):
        self._check(lib().gx_build_reset(self._h))

    def close(self):
        if self._h:
            lib().gx_build_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def unitig_limits() -> dict:
    """The unitig pass's limits (no device needed): distinct k-mers per
    graph and the last letter a read head may sit at (ReadHeadInfo's 23-bit
    offset field)."""
    n, o = ctypes.c_uint64(), ctypes.c_uint64()
    lib().gx_unitig_limits(ctypes.byref(n), ctypes.byref(o))
    return {"max_nodes": n.value, "max_head_offset": o.value}


def build_and_merge(text: bytes, k: int, edge_order: str = "jdk7", device: int = 0):
    """Graph build followed by unitig compression; returns (graph records,
    unitig records)."""
    with GraphBuilder(k, edge_order=edge_order, device=device) as b:
        b.add_reads(text)
        b.finish()
        g = b.part_bytes(0)
        return g, b.unitigs()


def build_graph(text: bytes, k: int, nparts: int = 1, edge_order: str = "jdk7", device: int = 0, method=None):
    """Graph build of one readids buffer; returns the per-part record streams."""
    with GraphBuilder(k, nparts=nparts, edge_order=edge_order, device=device, method=method) as b:
        b.add_reads(text)
        b.finish()
        return [b.part_bytes(p) for p in range(nparts)]


def fit_mixture(values, counts, max_value=None, iterations=10) -> dict:
    """FittingMixture.fittingMixture over binned data (gx_fit_mixture): value
    values[i] occurring counts[i] times; max_value defaults to the largest
    value.  Returns MIXTURE_KEYS -> value."""
    n = len(values)
    if len(counts) != n:
        raise ValueError("values and counts differ in length")
    v = (ctypes.c_double * max(n, 1))(*[float(x) for x in values])
    c = (ctypes.c_uint64 * max(n, 1))(*[int(x) for x in counts])
    mx = max((float(x) for x in values), default=0.0) if max_value is None else float(max_value)
    a = (ctypes.c_double * len(MIXTURE_KEYS))()
    rc = lib().gx_fit_mixture(v, c, n, mx, iterations, a)
    if rc:
        raise GxError(rc, lib().gx_last_error(None).decode())
    return dict(zip(MIXTURE_KEYS, list(a)))


STAT_KEYS = ["occurrences", "distinct", "heads", "table_capacity", "replayed_nodes", "replayed_occurrences",
             "max_coverage", "predicted_distinct", "table_rebuilds"]


class _View:
    """The build handle a batch consumer receives (valid during the callback)."""

    def __init__(self, h, nparts):
        self._h = ctypes.c_void_p(h)
        self.nparts = nparts

    _check = GraphBuilder._check
    part_info = GraphBuilder.part_info
    part_bytes = GraphBuilder.part_bytes
    stats = GraphBuilder.stats
    unitig_bytes = GraphBuilder.unitig_records


def load_tuning(path=os.path.join(HERE, "batch_tuning.json")):
    """The autotuner's committed choice (genomix_amd/autotune.py), or None."""
    if not os.path.exists(path):
        return None
    import json
    with open(path) as f:
        t = json.load(f)
    out = {k: t[k] for k in ("workers", "max_occ", "grid", "slack")}
    if "method" in t:
        out["method"] = t["method"]
    return out


class GraphBatch:
    """Many read sets (mixed k) built in one call by the native scheduler
    (include/genomix_batch.h).  Inputs: bytes (host) or (device pointer,
    length) pairs; they must stay alive until run() returns."""

    def __init__(self, sets, device=0, sizer=None, tuning=None):
        self._b = ctypes.c_void_p()
        rc = lib().gx_batch_create(device, len(sets), ctypes.byref(self._b))
        if rc != 0:
            raise GxError(rc, "gx_batch_create failed")
        self._keep = []
        self.sets = sets
        for i, s in enumerate(sets):
            cfg = _Cfg(s["k"], s.get("nparts", 1), 250, EDGE_ORDER[s.get("edge_order", "jdk7")], device, 0)
            data = s["input"]
            if isinstance(data, tuple):
                ptr, n, on_dev = ctypes.c_void_p(data[0]), data[1], 1
            else:
                buf = ctypes.create_string_buffer(data, len(data))
                self._keep.append(buf)
                ptr, n, on_dev = ctypes.cast(buf, ctypes.c_void_p), len(data), 0
            self._check(lib().gx_batch_set(self._b, i, ctypes.byref(cfg), ptr, n, s.get("fmt", GX_FMT_READIDS),
                                           on_dev))
        self._sizer = sizer
        if sizer is not None:
            self._check(lib().gx_batch_set_sizer(self._b, sizer.handle))
        if tuning is None:
            tuning = load_tuning()
        if tuning is not None:
            self.set_tuning(**tuning)

    def _check(self, rc):
        if rc != 0:
            raise GxError(rc, lib().gx_batch_last_error(self._b).decode(errors="replace"))

    def set_tuning(self, workers, max_occ, grid, slack, method=None):
        """Workers, size classes (max_occ bounds) and per class the grid,
        table slack and aggregation method ("sort" or "hash"; default sort)."""
        n = len(grid)
        mo = (ctypes.c_uint64 * max(1, n - 1))(*max_occ)
        g = (ctypes.c_int * n)(*grid)
        sl = (ctypes.c_float * n)(*slack)
        self._check(lib().gx_batch_set_tuning(self._b, workers, n, mo, g, sl))
        if method is not None:
            m = (ctypes.c_int * len(method))(*[METHOD[x] for x in method])
            self._check(lib().gx_batch_set_methods(self._b, len(method), m))

    def run(self, unitigs=True, collect=False, consumer=None):
        """Builds every set.  collect=True returns {set: (parts, unitig bytes)};
        consumer(set, view) is called on worker threads otherwise."""
        import threading
        out = {}
        lock = threading.Lock()
        errors = []

        def cb(user, si, h):
            try:
                v = _View(h, self.sets[si].get("nparts", 1))
                if collect:
                    parts = [v.part_bytes(p) for p in range(v.nparts)]
                    u = v.unitig_bytes() if unitigs else None
                    with lock:
                        out[si] = (parts, u)
                elif consumer is not None:
                    consumer(si, v)
                return 0
            except Exception as e:  # reported by the batch as a consumer failure
                errors.append(e)
                return -1

        fn = BATCH_CONSUMER(cb) if (collect or consumer is not None) else BATCH_CONSUMER()
        rc = lib().gx_batch_run(self._b, int(unitigs), fn, None)
        if errors:
            raise errors[0]
        self._check(rc)
        return out

    def retried(self):
        """Sets of the last run built again alone for want of device memory."""
        return lib().gx_batch_retried(self._b)

    def set_stats(self, i):
        a = (ctypes.c_uint64 * 9)()
        ms = ctypes.c_double()
        self._check(lib().gx_batch_set_stats(self._b, i, a, 9, ctypes.byref(ms)))
        d = dict(zip(STAT_KEYS, list(a)))
        d["ms"] = ms.value
        return d

    def close(self):
        if self._b:
            lib().gx_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
