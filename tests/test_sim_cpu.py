"""The synthetic read generator: the vectorised readids writer gives the
per-read writer's bytes, and the chunked generator numbers reads on across
chunks over the one genome."""

import numpy as np

from genomix_amd import sim


def test_readids_writers_agree():
    rng = np.random.default_rng(3)
    for n, L, first in [(1, 5, 0), (37, 12, 0), (200, 30, 2499), (5, 3, 10 ** 9)]:
        m1 = rng.integers(0, 4, (n, L), dtype=np.uint8)
        m2 = rng.integers(0, 4, (n, L), dtype=np.uint8)
        assert sim.readids_text_np(m1, None, first) == sim.readids_text(m1, None, first)
        assert sim.readids_text_np(m1, m2, first) == sim.readids_text(m1, m2, first)


def test_chunks_number_reads_on():
    spec = sim.ReadSetSpec(genome_len=5000, read_len=50, coverage=10, error_rate=0.01, paired=True,
                           insert_mean=150, insert_sd=10, repeat_frac=0.1, repeat_len=(100, 300), seed=8)
    chunks = list(sim.simulate_chunks(spec, chunk_reads=128))
    text = b"".join(chunks)
    lines = text.split(b"\n")[:-1]
    assert len(lines) == sim.n_reads(spec) == 500
    assert len(chunks) == 4
    assert [int(ln.split(b"\t")[0]) for ln in lines] == [4 * i + 2 for i in range(500)]
    assert all(len(ln.split(b"\t")[1]) == 50 == len(ln.split(b"\t")[2]) for ln in lines)
    assert list(sim.simulate_chunks(spec, chunk_reads=128)) == chunks  # seeded
