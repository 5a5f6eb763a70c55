"""Autotuner of the batch scheduler (include/genomix_batch.h).

Sweeps, on a sample of the batched workload, the knobs the runtime exposes:
per size class of read set (k-mer occurrences <= 2^23, <= 2^27, larger) the
aggregation method -- sorted, or the hash table sized by the learned
distinct-k-mer prediction with a slack (capacity / predicted distinct
k-mers: a longer or shorter probe sequence against a smaller or larger
table) -- and the grid of the grid-stride kernels, then the number of
worker streams.  Per-class
times come from the scheduler's own per-set clock with one worker; the
worker count from the wall time of whole-batch runs.  Writes
genomix_amd/batch_tuning.json, which bench.py and GraphBatch users load.

    python -m genomix_amd.autotune [--out DIR]
"""

import argparse
import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CLASSES = [1 << 23, 1 << 27]
GRIDS = [512, 1024, 2048, 4096]
SLACKS = [1.25, 1.5, 2.0]
WORKERS = [1, 2, 3, 4, 6, 8]
# (method, grid, slack); the slack only sizes the hash method's table
CANDIDATES = [("sort", g, 1.5) for g in GRIDS] + [("hash", g, sl) for g in GRIDS for sl in SLACKS]


def _class_of(occ):
    c = 0
    while c < len(CLASSES) and occ > CLASSES[c]:
        c += 1
    return c


def main():
    from . import devmem, gx, sim, sizer
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--sets", type=int, default=16)
    ap.add_argument("--scale", type=float, default=0.5)
    args = ap.parse_args()

    def log(m):
        print(m, flush=True)

    specs = sim.batched_sets(n=args.sets, seed=777, scale=args.scale)
    texts = sim.simulate_many([sp for sp, _ in specs])
    dev = [devmem.DeviceBytes(t) for t in texts]
    sets = [{"k": k, "input": (d.address, d.n)} for (sp, k), d in zip(specs, dev)]
    log(f"sample: {len(sets)} sets, {sum(map(len, texts)) / 1e9:.2f} GB of reads")
    del texts
    net = sizer.load()
    per_class = {c: {} for c in range(len(CLASSES) + 1)}
    with gx.GraphBatch(sets, sizer=net) as b:
        occ = None
        for cand in CANDIDATES:
            method, grid, slack = cand
            b.set_tuning(1, CLASSES, [grid] * 3, [slack] * 3, method=[method] * 3)
            b.run(unitigs=True)  # warm
            b.run(unitigs=True)
            st = [b.set_stats(i) for i in range(len(sets))]
            occ = [x["occurrences"] for x in st]
            for c in per_class:
                per_class[c][cand] = sum(x["ms"] for x, o in zip(st, occ) if _class_of(o) == c)
            log(f"{method} grid {grid} slack {slack}: " +
                ", ".join(f"class {c} {per_class[c][cand]:.1f} ms" for c in per_class))
        best = {}
        for c, d in per_class.items():
            if not any(v > 0 for v in d.values()):
                best[c] = ("sort", 2048, 1.5)  # no sample set in this class
            else:
                best[c] = min(d, key=d.get)
        method = [best[c][0] for c in sorted(best)]
        grid = [best[c][1] for c in sorted(best)]
        slack = [best[c][2] for c in sorted(best)]
        wall = {}
        for w in WORKERS:
            b.set_tuning(w, CLASSES, grid, slack, method=method)
            b.run(unitigs=True)
            ts = []
            for _ in range(3):
                devmem.synchronize()
                t0 = time.perf_counter()
                b.run(unitigs=True)
                ts.append(time.perf_counter() - t0)
            wall[w] = float(np.median(ts))
            log(f"workers {w}: {wall[w] * 1e3:.1f} ms per batch")
    for d in dev:
        d.free()
    workers = min(wall, key=wall.get)
    out = {"workers": workers, "max_occ": CLASSES, "grid": grid, "slack": slack, "method": method,
           "measured": {"per_class_ms": {str(c): {f"{m}/{g}/{sl}": round(v, 3) for (m, g, sl), v in d.items()}
                                         for c, d in per_class.items()},
                        "batch_ms_by_workers": {str(w): round(v * 1e3, 3) for w, v in wall.items()}},
           "sample": f"{len(sets)} sets of the batched workload (seed 777) at genome scale {args.scale}",
           "class_sets": {str(c): sum(1 for o in occ if _class_of(o) == c) for c in per_class}}
    with open(os.path.join(args.out, "batch_tuning.json"), "w") as f:
        json.dump(out, f, indent=1)
    log(json.dumps(out))


if __name__ == "__main__":
    main()
