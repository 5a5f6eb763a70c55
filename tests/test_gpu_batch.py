"""The batched runtime (include/genomix_batch.h): many read sets of uneven
size and mixed k in one call give, set by set, the bytes of separate builds
and the oracle's graph and unitigs."""

import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gx():
    import genomix_amd.gx as g
    return g


def _sets():
    from genomix_amd import sim
    out = []
    specs = [(3000, 70, 10, False, 21), (12000, 100, 15, True, 31), (1500, 60, 6, False, 55),
             (20000, 150, 12, True, 55), (5000, 100, 20, False, 31), (800, 50, 5, True, 21),
             (9000, 120, 8, True, 21), (2500, 90, 25, False, 33)]
    for i, (G, L, cov, paired, k) in enumerate(specs):
        spec = sim.ReadSetSpec(genome_len=G, read_len=L, coverage=cov, error_rate=0.01, paired=paired,
                               insert_mean=3 * L, insert_sd=L // 4, repeat_frac=0.05 if G > 5000 else 0.0,
                               repeat_len=(300, 900), repeat_div=0.01, seed=4242 + i)
        out.append({"k": k, "nparts": 1 + (i % 3), "input": sim.simulate(spec)})
    return out


@pytest.mark.parametrize("workers,method", [(1, ["sort"] * 3), (4, ["hash", "sort", "hash"]), (2, ["hash"] * 3)])
def test_batch_matches_oracle_per_set(gx, workers, method):
    """Every set equals the oracle whichever aggregation method its size
    class is given (a worker's handle switches method between sets)."""
    sets = _sets()
    tuning = {"workers": workers, "max_occ": [20000, 200000], "grid": [256, 1024, 2048],
              "slack": [1.5, 1.5, 2.0], "method": method}
    with gx.GraphBatch(sets, tuning=tuning) as b:
        out = b.run(unitigs=True, collect=True)
        assert sorted(out) == list(range(len(sets)))
        assert b.retried() == 0  # small sets: none runs out of memory beside the others
        for i, s in enumerate(sets):
            parts, u = out[i]
            want = oracle.build_graph(s["input"], s["k"], nparts=s["nparts"])
            assert parts == want, f"set {i}"
            assert u == oracle.unitigs(oracle.build_graph(s["input"], s["k"])[0], s["k"]), f"set {i} unitigs"
            st = b.set_stats(i)
            assert st["occurrences"] > 0 and st["ms"] > 0


def test_batch_with_learned_sizing_and_device_inputs(gx):
    import os
    from genomix_amd import sizer
    from genomix_amd.devmem import DeviceBytes
    if not os.path.exists(sizer.WEIGHTS):
        pytest.skip("sizing weights not trained yet")
    sets = _sets()
    dev = [DeviceBytes(s["input"]) for s in sets]
    dsets = [dict(s, input=(d.address, d.n)) for s, d in zip(sets, dev)]
    net = sizer.load()
    # every class on the hash method: the network sizes each set's table
    tuning = {"workers": 3, "max_occ": [20000, 200000], "grid": [1024] * 3, "slack": [1.5] * 3,
              "method": ["hash"] * 3}
    try:
        with gx.GraphBatch(dsets, sizer=net, tuning=tuning) as b:
            out = b.run(unitigs=False, collect=True)
            for i, s in enumerate(sets):
                assert out[i][0] == oracle.build_graph(s["input"], s["k"], nparts=s["nparts"])
                st = b.set_stats(i)
                assert st["predicted_distinct"] > 0
                # ceil(1.5 x prediction) slots (at least 1024), or more after a rebuild
                assert st["table_capacity"] >= max(1024, int(1.5 * st["predicted_distinct"]))
    finally:
        for d in dev:
            d.free()
    # twice in a row on the same batch (worker handles are reused)
    with gx.GraphBatch(sets) as b:
        a1 = b.run(unitigs=True, collect=True)
        a2 = b.run(unitigs=True, collect=True)
        assert a1 == a2


def test_batch_method_per_class_is_checked(gx):
    sets = _sets()[:2]
    with gx.GraphBatch(sets, tuning={"workers": 1, "max_occ": [20000], "grid": [512, 512], "slack": [1.5, 1.5]}) as b:
        with pytest.raises(gx.GxError) as e:  # three methods for two classes
            b.set_tuning(1, [20000], [512, 512], [1.5, 1.5], method=["sort", "hash", "hash"])
        assert e.value.code == gx.GX_EINVAL
        m = (gx.ctypes.c_int * 2)(0, 7)
        assert gx.lib().gx_batch_set_methods(b._b, 2, m) == gx.GX_EINVAL  # unknown method


def test_batch_reports_the_failing_set(gx):
    sets = _sets()[:3]
    sets.append({"k": 21, "input": b"2\tACGTACGTACGTACGTACGTACGT\nnot-a-number\tACGT\n"})
    with gx.GraphBatch(sets, tuning={"workers": 2, "max_occ": [], "grid": [1024], "slack": [1.5]}) as b:
        with pytest.raises(gx.GxError) as e:
            b.run(unitigs=False)
        assert e.value.code == gx.GX_EINPUT
        assert "set 3" in str(e.value)
