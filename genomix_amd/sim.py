"""Seeded synthetic genomes and reads (readids text), vectorised with numpy.

Genome: iid uniform ACGT; optionally repeat families: copies of random
templates, each copy with iid substitutions at the stated divergence, placed
until the stated fraction of the genome is covered (one template of a fixed
length, or for a length range, families of 2-10 copies with lengths drawn
from the range).  Reads: start uniform,
strand uniform, iid substitution errors (one of the other three bases).
Paired-end: fragment length ~ round(N(mean, sd)) clipped to [L, G]; mate 1 is
the fragment's first L bases, mate 2 the reverse complement of its last L;
on the reverse strand the fragment is reverse-complemented first (the same
pairing as genomix-driver/src/test/python/generate_graph.py).  Lines follow
the FASTQ conversion of GenomixDriver.convertAndUploadFastqToHDFS
(GenomixDriver.java:665-714): readId = 4i + 2.
"""

from dataclasses import dataclass

import numpy as np

ASCII = np.frombuffer(b"ACGT", dtype=np.uint8)


@dataclass
class ReadSetSpec:
    genome_len: int
    read_len: int
    coverage: float
    error_rate: float
    paired: bool = False
    insert_mean: int = 500
    insert_sd: int = 50
    repeat_frac: float = 0.0
    repeat_len: int = 2000          # or (lo, hi) for a family of lengths
    repeat_div: float = 0.0
    seed: int = 20261202


def genome(spec: ReadSetSpec, rng) -> np.ndarray:
    g = rng.integers(0, 4, spec.genome_len, dtype=np.uint8)
    if spec.repeat_frac > 0:
        covered = 0
        target = int(spec.repeat_frac * spec.genome_len)
        fam = None
        left = 0  # copies left of the current family (length ranges: 2..10 copies per family)
        while covered < target:
            if isinstance(spec.repeat_len, tuple):
                if left == 0:
                    ln = int(rng.integers(spec.repeat_len[0], spec.repeat_len[1] + 1))
                    fam = rng.integers(0, 4, ln, dtype=np.uint8)
                    left = int(rng.integers(2, 11))
                left -= 1
            elif fam is None:
                fam = rng.integers(0, 4, spec.repeat_len, dtype=np.uint8)
            ln = min(len(fam), spec.genome_len)
            copy = fam[:ln].copy()
            mut = rng.random(ln) < spec.repeat_div
            copy[mut] = (copy[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
            pos = int(rng.integers(0, spec.genome_len - ln + 1))
            g[pos:pos + ln] = copy
            covered += ln
    return g


def _add_errors(reads, rate, rng):
    if rate <= 0:
        return reads
    mut = rng.random(reads.shape) < rate
    reads[mut] = (reads[mut] + rng.integers(1, 4, int(mut.sum()), dtype=np.uint8)) % 4
    return reads


def _rc(codes):
    return (3 - codes[..., ::-1]).astype(np.uint8)


def n_reads(spec: ReadSetSpec, G=None):
    """Reads (single-end) or pairs (paired-end) of a read set."""
    G = spec.genome_len if G is None else G
    return int(round(spec.coverage * G / ((2 if spec.paired else 1) * spec.read_len)))


def reads(spec: ReadSetSpec, g=None, rng=None, n=None):
    """Returns (mate1, mate2 or None) code matrices of shape (n, L); n
    defaults to the spec's coverage."""
    rng = rng if rng is not None else np.random.default_rng(spec.seed)
    if g is None:
        g = genome(spec, rng)
    G, L = len(g), spec.read_len
    if n is None:
        n = n_reads(spec, G)
    if spec.paired:
        frag = np.rint(rng.normal(spec.insert_mean, spec.insert_sd, n)).astype(np.int64)
        frag = np.clip(frag, L, G)
        start = (rng.random(n) * (G - frag + 1)).astype(np.int64)
        rev = rng.random(n) < 0.5
        win = np.lib.stride_tricks.sliding_window_view(g, L)  # row s = g[s:s + L]
        a = win[start]
        b = win[start + frag - L]
        # forward: (start, rc(end)); reverse strand = the reverse-complemented
        # fragment: (rc(end), rc(rc(start)) = start)
        m1 = np.where(rev[:, None], _rc(b), a)
        m2 = np.where(rev[:, None], a, _rc(b))
        return _add_errors(m1.astype(np.uint8), spec.error_rate, rng), _add_errors(
            m2.astype(np.uint8), spec.error_rate, rng)
    start = (rng.random(n) * (G - L + 1)).astype(np.int64)
    rev = rng.random(n) < 0.5
    r = np.lib.stride_tricks.sliding_window_view(g, L)[start]
    r = np.where(rev[:, None], _rc(r), r)
    return _add_errors(r.astype(np.uint8), spec.error_rate, rng), None


def readids_text(m1, m2=None, first_read=0) -> bytes:
    """Format code matrices as readids lines (readId = 4i + 2)."""
    n, L = m1.shape
    ids = [b"%d\t" % (4 * (first_read + i) + 2) for i in range(n)]
    s1 = ASCII[m1]
    if m2 is None:
        body = np.concatenate([s1, np.full((n, 1), ord("\n"), np.uint8)], axis=1)
    else:
        s2 = ASCII[m2]
        body = np.concatenate([s1, np.full((n, 1), ord("\t"), np.uint8), s2,
                               np.full((n, 1), ord("\n"), np.uint8)], axis=1)
    rows = body.tobytes()
    w = body.shape[1]
    return b"".join(ids[i] + rows[i * w:(i + 1) * w] for i in range(n))


def readids_text_np(m1, m2=None, first_read=0) -> bytes:
    """readids_text() with no per-read Python work (the same bytes): the
    lines are laid out in a byte matrix whose id columns are left-padded
    with zero bytes, which are then dropped."""
    n, L = m1.shape
    x = 4 * (first_read + np.arange(n, dtype=np.int64)) + 2
    D = len(str(int(x[-1]))) if n else 1
    dig = np.zeros((n, D), np.uint8)
    for c in range(D - 1, -1, -1):
        dig[:, c] = np.where((x > 0) | (c == D - 1), 48 + x % 10, 0)
        x //= 10
    cols = [dig, np.full((n, 1), ord("\t"), np.uint8), ASCII[m1]]
    if m2 is not None:
        cols += [np.full((n, 1), ord("\t"), np.uint8), ASCII[m2]]
    cols.append(np.full((n, 1), ord("\n"), np.uint8))
    flat = np.concatenate(cols, axis=1).ravel()
    return flat[flat != 0].tobytes()


_chunk_genome = {}


def _chunk_text(args):
    spec, c, first, n = args
    g = _chunk_genome.get(spec.seed)
    if g is None:
        _chunk_genome.clear()
        g = _chunk_genome[spec.seed] = genome(spec, np.random.default_rng(spec.seed))
    m1, m2 = reads(spec, g, np.random.default_rng([spec.seed, c + 1]), n)
    return readids_text_np(m1, m2, first_read=first)


def simulate_chunks(spec: ReadSetSpec, chunk_reads=1 << 20, workers=1):
    """A read set too large for one array, as readids text chunks of
    `chunk_reads` reads (pairs) each, in order, numbered on across chunks.
    The genome is the one simulate() draws for the spec; chunk c's reads
    come from a generator seeded with (seed, c + 1), so the chunks are fixed
    by the seed and the chunk size and can be drawn by `workers` processes
    (each draws the genome once)."""
    total = n_reads(spec)
    jobs = [(spec, c, s, min(chunk_reads, total - s)) for c, s in enumerate(range(0, total, chunk_reads))]
    if workers <= 1:
        for j in jobs:
            yield _chunk_text(j)
        return
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(workers) as pool:
        for t in pool.imap(_chunk_text, jobs):
            yield t
        # let the workers exit on their own: the context manager's terminate()
        # would signal them (a profiler's preload then reports an abort)
        pool.close()
        pool.join()


def fastq_text(m, name="r") -> bytes:
    """Code matrix as FASTQ records (quality 'I')."""
    n, L = m.shape
    seqs = ASCII[m].tobytes()
    qual = b"I" * L
    return b"".join(b"@%s%d\n%s\n+\n%s\n" % (name.encode(), i, seqs[i * L:(i + 1) * L], qual) for i in range(n))


def simulate(spec: ReadSetSpec) -> bytes:
    m1, m2 = reads(spec)
    return readids_text(m1, m2)


# Workloads named by what they are (sizes as in the benchmark plan).
PLAIN_CPU_50KBP = ReadSetSpec(genome_len=50_000, read_len=100, coverage=30, error_rate=0.005, seed=20261203)
PLAIN_GPU_5MBP = ReadSetSpec(genome_len=5_000_000, read_len=150, coverage=50, error_rate=0.01, paired=True,
                             insert_mean=500, insert_sd=50, repeat_frac=0.05, repeat_len=2000, repeat_div=0.01,
                             seed=20261204)


# Learned sizing applied: 100 Mbp genome, 20% repeat families of 1-10 kb at
# 2% divergence, 150 bp PE, 40x, 1% error (built with k = 31).
LEARNED_100MBP = ReadSetSpec(genome_len=100_000_000, read_len=150, coverage=40, error_rate=0.01, paired=True,
                             insert_mean=500, insert_sd=50, repeat_frac=0.2, repeat_len=(1000, 10000),
                             repeat_div=0.02, seed=20261207)


# Large streamed: 500 Mbp genome, 30% repeat families of 1-10 kb at 2%
# divergence, 150 bp PE (insert 400/40), 30x, 1% error (built with k = 55).
STREAMED_500MBP = ReadSetSpec(genome_len=500_000_000, read_len=150, coverage=30, error_rate=0.01, paired=True,
                              insert_mean=400, insert_sd=40, repeat_frac=0.3, repeat_len=(1000, 10000),
                              repeat_div=0.02, seed=20261208)


def batched_sets(n=64, seed=20261206, scale=1.0):
    """The batched workload: n independent read sets, genome sizes
    log-uniform in [100 kbp, 10 Mbp] (times `scale`), 100 or 150 bp reads,
    single or paired, 20-80x, 0.5-2% error, k drawn from 21/31/55.
    Returns [(spec, k)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        G = int(np.exp(rng.uniform(np.log(1e5), np.log(1e7))) * scale)
        L = int(rng.choice([100, 150]))
        paired = bool(rng.random() < 0.5)
        spec = ReadSetSpec(genome_len=max(G, 4 * L), read_len=L, coverage=float(rng.uniform(20, 80)),
                           error_rate=float(rng.uniform(0.005, 0.02)), paired=paired,
                           insert_mean=int(3.3 * L), insert_sd=L // 3, seed=seed + 1000 + i)
        out.append((spec, int(rng.choice([21, 31, 55]))))
    return out


def _simulate_indexed(args):
    i, spec = args
    return i, simulate(spec)


def simulate_many(specs, workers=12):
    """simulate() of each spec, in worker processes; returns texts in order."""
    import multiprocessing as mp
    out = [None] * len(specs)
    if workers <= 1:
        return [simulate(s) for s in specs]
    with mp.get_context("spawn").Pool(workers) as pool:
        for i, t in pool.imap_unordered(_simulate_indexed, list(enumerate(specs))):
            out[i] = t
        pool.close()
        pool.join()
    return out
