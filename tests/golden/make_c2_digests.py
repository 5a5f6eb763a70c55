"""Writes tests/golden/digests/plain_gpu_params.json: SHA-256 digests of the
oracle's graph records (part-0, P = 1) and unitig records for read sets with
the parameters of the 5 Mbp plain GPU workload (5% of the genome in 2 kb
repeat copies at 1% divergence, 150 bp paired-end reads, insert 500/50, 50x,
1% substitution errors, k = 31) at smaller genome sizes.  The oracle takes
minutes at these sizes, so the GPU test compares against these digests
instead of running it.

Run from the repository root:  python tests/golden/make_c2_digests.py
"""

import dataclasses
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from genomix_amd import sim  # noqa: E402
from oracle import oracle  # noqa: E402

SIZES = (500_000, 1_000_000)
OUT = os.path.join(ROOT, "tests", "golden", "digests", "plain_gpu_params.json")


def spec_for(G):
    return dataclasses.replace(sim.PLAIN_GPU_5MBP, genome_len=G, seed=20261204 + G)


def main():
    cases = []
    for G in SIZES:
        spec = spec_for(G)
        text = sim.simulate(spec)
        t0 = time.time()
        graph = oracle.build_graph(text, 31)[0]
        t1 = time.time()
        merged = oracle.unitigs(graph, 31)
        t2 = time.time()
        cases.append({
            "genome_len": G, "seed": spec.seed, "k": 31,
            "text_sha256": hashlib.sha256(text).hexdigest(), "text_bytes": len(text),
            "graph_sha256": hashlib.sha256(graph).hexdigest(), "graph_bytes": len(graph),
            "unitigs_sha256": hashlib.sha256(merged).hexdigest(), "unitigs_bytes": len(merged),
            "oracle_seconds": {"graph": round(t1 - t0, 1), "unitigs": round(t2 - t1, 1)},
        })
        print(cases[-1], flush=True)
    with open(OUT, "w") as f:
        json.dump({"workload": "plain GPU workload parameters at smaller genome sizes", "cases": cases}, f,
                  indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
