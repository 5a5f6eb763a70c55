"""Learned table sizing: the FP32 MLP on the device and its training.

The network (include/genomix_sizer.h, genomix_amd/csrc/gx_sizer.hip) maps
the 72 sizing features that the graph build computes at finish (a 2^12-
register HyperLogLog sketch of the first 1% of the mates, k, read length,
a coverage estimate, the partition count) to

    y0 = log(distinct k-mers / k-mer occurrences)
    y1 = log(largest partition's share of the nodes * P)

Training targets are exact: every training read set is built by the graph
build itself (``distinct`` from its statistics, partition sizes from its
parts).  The forward, backward and Adam kernels are the project's own; the
gradients are checked against a float64 CPU restatement in the tests.

``python -m genomix_amd.sizer --runs 10000`` regenerates the committed
weights (``sizer_mlp.npy``) on a GPU and writes a training summary
(``sizer_mlp.json``).
"""

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import gx

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "sizer_mlp.npy")
SUMMARY = os.path.join(HERE, "sizer_mlp.json")
DIN, HIDDEN, DOUT = 72, 64, 2

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


class Mlp:
    """One gx_mlp handle (parameters and Adam moments in HBM)."""

    def __init__(self, din=DIN, hidden=HIDDEN, dout=DOUT, device=0, params=None):
        self.handle = ctypes.c_void_p()
        rc = gx.lib().gx_mlp_create(din, hidden, dout, device, ctypes.byref(self.handle))
        if rc != 0:
            raise gx.GxError(rc, "gx_mlp_create failed")
        self.din, self.hidden, self.dout = din, hidden, dout
        self.n = gx.lib().gx_mlp_num_params(self.handle)
        if params is not None:
            self.set_params(params)

    def _check(self, rc):
        if rc != 0:
            raise gx.GxError(rc, gx.lib().gx_mlp_last_error(self.handle).decode())

    def set_params(self, p):
        a, ptr = _f32(p)
        self._check(gx.lib().gx_mlp_set_params(self.handle, ptr, a.size))

    def params(self):
        a = np.zeros(self.n, np.float32)
        self._check(gx.lib().gx_mlp_get_params(self.handle, a.ctypes.data_as(_fp), self.n))
        return a

    def forward(self, x):
        x, xp = _f32(x)
        y = np.zeros((x.shape[0], self.dout), np.float32)
        self._check(gx.lib().gx_mlp_forward(self.handle, xp, x.shape[0], y.ctypes.data_as(_fp)))
        return y

    def gradient(self, x, t):
        x, xp = _f32(x)
        t, tp = _f32(t)
        g = np.zeros(self.n, np.float32)
        loss = ctypes.c_float()
        self._check(gx.lib().gx_mlp_gradient(self.handle, xp, tp, x.shape[0], g.ctypes.data_as(_fp),
                                             ctypes.byref(loss)))
        return g, loss.value

    def train_step(self, x, t, lr):
        x, xp = _f32(x)
        t, tp = _f32(t)
        loss = ctypes.c_float()
        self._check(gx.lib().gx_mlp_train_step(self.handle, xp, tp, x.shape[0], ctypes.c_float(lr),
                                               ctypes.byref(loss)))
        return loss.value

    def close(self):
        if self.handle:
            gx.lib().gx_mlp_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_params(din=DIN, hidden=HIDDEN, dout=DOUT, seed=0):
    """Glorot-uniform weights, zero biases, in the flat layout W1 b1 W2 b2 W3 b3."""
    rng = np.random.default_rng(seed)
    out = []
    for fi, fo in ((din, hidden), (hidden, hidden), (hidden, dout)):
        lim = math.sqrt(6.0 / (fi + fo))
        out.append(rng.uniform(-lim, lim, fo * fi))
        out.append(np.zeros(fo))
    return np.concatenate(out).astype(np.float32)


def load(device=0, path=WEIGHTS):
    """The committed trained network."""
    p = np.load(path, allow_pickle=False)
    return Mlp(device=device, params=p)


def train(mlp, X, T, epochs=300, batch=256, lr=1e-3, seed=0, log=None):
    """Minibatch Adam on the device; returns the loss per epoch."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    hist = []
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        for s in range(0, n, batch):
            idx = order[s:s + batch]
            tot += mlp.train_step(X[idx], T[idx], lr) * len(idx)
        hist.append(tot / n)
        if log and (ep % 25 == 0 or ep == epochs - 1):
            log(f"epoch {ep}: loss {hist[-1]:.5f}")
    return hist


# ---------------------------------------------------------------------------
# training data: read sets built by the graph build itself
# ---------------------------------------------------------------------------

def random_spec(rng, seed):
    from .sim import ReadSetSpec
    G = int(math.exp(rng.uniform(math.log(2e4), math.log(1e6))))
    L = int(rng.choice([100, 150]))
    paired = bool(rng.random() < 0.5)
    rl = (1000, max(1000, min(10000, G // 4)))
    return ReadSetSpec(genome_len=G, read_len=L, coverage=float(rng.uniform(5, 80)),
                       error_rate=float(rng.uniform(0.005, 0.02)), paired=paired,
                       insert_mean=int(rng.uniform(2.5, 4) * L), insert_sd=int(L // 4),
                       repeat_frac=float(rng.uniform(0, 0.3)), repeat_len=rl,
                       repeat_div=float(rng.uniform(0, 0.03)), seed=seed)


def _make_text(args):
    i, base_seed = args
    from . import sim
    rng = np.random.default_rng(base_seed + i)
    spec = random_spec(rng, base_seed + 7919 * i)
    k = int(rng.choice([21, 31, 55]))
    P = int(rng.choice([1, 2, 4, 8, 16, 32, 64]))
    return i, k, P, sim.simulate(spec)


def targets_of(builder, P):
    st = builder.stats()
    D = st["distinct"]
    parts = [builder.part_info(p)[0] for p in range(P)]
    skew = max(parts) * P / max(D, 1)
    return st["occurrences"], D, skew


def generate(runs, base_seed=20261205, workers=12, log=None):
    """Features and exact targets of `runs` seeded read sets."""
    import multiprocessing as mp
    X = np.zeros((runs, DIN), np.float32)
    T = np.zeros((runs, DOUT), np.float32)
    meta = np.zeros((runs, 5), np.float64)  # k, P, occurrences, distinct, skew
    t0 = time.time()
    ctx = mp.get_context("spawn")
    builders = {}  # one handle per (k, P), reused: device buffers stay allocated
    with ctx.Pool(workers) as pool:
        for n_done, (i, k, P, text) in enumerate(pool.imap_unordered(_make_text, [(i, base_seed) for i in
                                                                              range(runs)], chunksize=4)):
            b = builders.get((k, P))
            if b is None:
                b = builders[(k, P)] = gx.GraphBuilder(k, nparts=P)
            b.reset()
            b.add_reads(text)
            b.finish()
            f, _, _ = b.features()
            N, D, skew = targets_of(b, P)
            X[i] = f
            T[i] = [math.log(D / N), math.log(skew)]
            meta[i] = [k, P, N, D, skew]
            if log and (n_done + 1) % 500 == 0:
                log(f"{n_done + 1}/{runs} read sets built ({time.time() - t0:.0f} s)")
        pool.close()  # workers exit on their own rather than by terminate()
        pool.join()
    for b in builders.values():
        b.close()
    return X, T, meta


def evaluate(pred_y0, T, meta):
    """Relative error of the predicted distinct count and the share of sets
    whose table (1.5 x prediction) would overflow."""
    D = meta[:, 3]
    N = meta[:, 2]
    Dhat = N * np.exp(np.minimum(pred_y0, 0))
    rel = Dhat / D - 1
    return {"median_abs_rel_err": float(np.median(np.abs(rel))), "p90_abs_rel_err": float(np.quantile(np.abs(rel), 0.9)),
            "overflow_share": float(np.mean(1.5 * Dhat < D)),
            "mean_table_slots_over_distinct": float(np.mean(np.maximum(1024, np.ceil(1.5 * Dhat)) / D)),
            "mean_default_slots_over_distinct": float(np.mean([2 ** math.ceil(math.log2(max(n, 1024))) / d
                                                               for n, d in zip(N, D)]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--data-out", default=None, help="also save the training set (.npz)")
    ap.add_argument("--out-dir", default=HERE, help="where sizer_mlp.npy / sizer_mlp.json go")
    args = ap.parse_args()
    weights = os.path.join(args.out_dir, os.path.basename(WEIGHTS))
    summary_path = os.path.join(args.out_dir, os.path.basename(SUMMARY))

    def log(m):
        print(m, flush=True)

    X, T, meta = generate(args.runs, workers=args.workers, log=log)
    if args.data_out:
        np.savez(args.data_out, X=X, T=T, meta=meta)
    n_val = max(1, args.runs // 10)
    Xtr, Ttr, Xva, Tva, Mva = X[n_val:], T[n_val:], X[:n_val], T[:n_val], meta[:n_val]
    mlp = Mlp(params=init_params(seed=1))
    t0 = time.time()
    hist = train(mlp, Xtr, Ttr, epochs=args.epochs, log=log)
    log(f"trained in {time.time() - t0:.1f} s")
    pred = mlp.forward(Xva)
    ev = evaluate(pred[:, 0], Tva, Mva)
    ev_skew = float(np.median(np.abs(np.exp(pred[:, 1]) - np.exp(Tva[:, 1]))))
    np.save(weights, mlp.params())
    summary = {"runs": args.runs, "validation_runs": n_val, "epochs": args.epochs, "batch": 256, "lr": 1e-3,
               "final_train_loss": hist[-1], "validation": ev, "validation_median_abs_skew_err": ev_skew,
               "spec_ranges": "G log-uniform 20-1000 kbp; L 100/150; SE/PE; 5-80x; error 0.5-2%; "
                              "repeats 0-30% in families of 2-10 copies, 1-10 kb, 0-3% divergence; "
                              "k 21/31/55; P 1..64",
               "base_seed": 20261205}
    with open(summary_path, "w") as f:
        json.dump(summary, f, indent=1)
    log(json.dumps(summary))


if __name__ == "__main__":
    sys.exit(main())
