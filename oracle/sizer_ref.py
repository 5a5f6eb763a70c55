"""CPU restatement of the learned-sizing arithmetic (test infrastructure only).

Imported only by tests/ as the checker of the device kernels in
genomix_amd/csrc/gx_sizer.hip (MLP forward, backward, Adam) and of the HLL
sketch and sizing features in genomix_amd/csrc/gx_build.hip.  Nothing here
has a counterpart in the reference (Genomix never sizes its aggregation in
advance).  What pins it:

* MLP gradient: central finite differences in float64
  (tests/test_sizer_cpu.py), so a dropped term, a wrong sign or a transposed
  weight fails;
* HLL estimator: the estimate of n distinct hashed values stays within a few
  standard errors (1.04/64) of n from 10^2 to 10^6 values, across the
  linear-counting switch (Flajolet et al. 2007);
* k-mer hashing: the canonical k-mer rule and 2-bit code are the oracle's
  own (checked against the reference's vectors in test_oracle_primitives).
"""

import math

import numpy as np

M64 = (1 << 64) - 1
HLL_BITS = 12
HLL_M = 1 << HLL_BITS


def mix64(x):
    """SplitMix64 finaliser, as the device's mix64."""
    x &= M64
    x ^= x >> 30
    x = (x * 0xbf58476d1ce4e5b9) & M64
    x ^= x >> 27
    x = (x * 0x94d049bb133111eb) & M64
    x ^= x >> 31
    return x


def khash(v, k):
    if k <= 31:
        return mix64(v)
    lo, hi = v & M64, v >> 64
    return mix64(lo ^ mix64(hi + 0x9e3779b97f4a7c15))


def hll_hash(v, k):
    return mix64(khash(v, k) + 0x632be59bd9b4e019)


CODE = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3, ord("a"): 0, ord("c"): 1, ord("g"): 2, ord("t"): 3}


def canonical_kmers(seq: bytes, k):
    """V = sum code(s_i) 4^i over the k-mer, canonical = min(V, revcomp)."""
    out = []
    for p in range(len(seq) - k + 1):
        v = r = 0
        for i in range(k):
            v |= CODE[seq[p + i]] << (2 * i)
            r |= (3 - CODE[seq[p + k - 1 - i]]) << (2 * i)
        out.append(min(v, r))
    return out


def _valid(m, k):
    return len(m) >= k and all(c in b"ACGTacgt" for c in m)


def mates_of(text: bytes):
    """Mate sequences in mate-index order (two per line; absent = b"")."""
    mates = []
    for line in text.split(b"\n"):
        if not line:
            continue
        f = line.rstrip(b"\r").split(b"\t")
        mates.append(f[1] if len(f) > 1 else b"")
        mates.append(f[2] if len(f) > 2 else b"")
    return mates


def hll_registers(text: bytes, k):
    """Registers of the sizing sample: the first max(ceil(1%), 512) mates."""
    mates = mates_of(text)
    nm = len(mates)
    ns = min(nm, max((nm + 99) // 100, 512))
    regs = np.zeros(HLL_M, np.uint8)
    occ = valid = 0
    for m in mates[:ns]:
        if not _valid(m, k):
            continue
        km = canonical_kmers(m, k)
        occ += len(km)
        valid += 1
        for v in km:
            h = hll_hash(v, k)
            idx = h >> (64 - HLL_BITS)
            rest = ((h << HLL_BITS) & M64) | (1 << (HLL_BITS - 1))
            rank = 64 - rest.bit_length() + 1
            regs[idx] = max(regs[idx], rank)
    return regs, occ, valid, (ns / nm if nm else 0.0)


def hll_estimate(regs):
    m = HLL_M
    inv = 0.0
    for r in regs:
        inv += math.ldexp(1.0, -int(r))
    E = 0.7213 / (1 + 1.079 / m) * m * m / inv
    zeros = int(np.sum(regs == 0))
    if E <= 2.5 * m and zeros > 0:
        E = m * math.log(m / zeros)
    return E


def features(regs, occ_s, valid_s, frac, k, P):
    """The 72 sizing features (same layout and float32 rounding as the device)."""
    m = HLL_M
    f = np.zeros(72, np.float32)
    hist = np.bincount(np.minimum(regs, 63).astype(np.int64), minlength=64).astype(np.float64)
    E = hll_estimate(regs)
    center = int(math.floor(float(np.sum(regs.astype(np.float64))) / m + 0.5))
    for v in range(64):
        b = min(63, max(0, v - center + 32))
        f[b] = np.float32(f[b] + np.float32(hist[v] / m))
    ratio = min(2.0, E / occ_s) if occ_s > 0 else 1.0
    cov = 2.0 * max(0.0, 1.0 - ratio) / frac if frac > 0 else 0.0
    f[64] = ratio
    f[65] = math.log1p(cov) / 5.0
    f[66] = k / 64.0
    f[67] = ((occ_s / valid_s + k - 1) if valid_s > 0 else 0.0) / 256.0
    f[68] = frac * 10.0
    f[69] = math.log2(P) / 16.0
    f[70] = 0.0
    f[71] = 1.0
    return f


# ---------------------------------------------------------------------------
# MLP: x -> tanh(W1 x + b1) -> tanh(W2 a1 + b2) -> W3 a2 + b3, float64
# ---------------------------------------------------------------------------

def unpack(p, din, h, dout):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for fi, fo in ((din, h), (h, h), (h, dout)):
        W = p[o:o + fo * fi].reshape(fo, fi)
        o += fo * fi
        b = p[o:o + fo]
        o += fo
        out += [W, b]
    assert o == p.size
    return out


def forward(p, X, din, h, dout):
    W1, b1, W2, b2, W3, b3 = unpack(p, din, h, dout)
    X = np.asarray(X, np.float64)
    a1 = np.tanh(X @ W1.T + b1)
    a2 = np.tanh(a1 @ W2.T + b2)
    return a2 @ W3.T + b3, (a1, a2)


def loss(p, X, T, din, h, dout):
    Y, _ = forward(p, X, din, h, dout)
    return 0.5 * np.sum((Y - T) ** 2) / X.shape[0]


def gradient(p, X, T, din, h, dout):
    W1, b1, W2, b2, W3, b3 = unpack(p, din, h, dout)
    X = np.asarray(X, np.float64)
    B = X.shape[0]
    Y, (a1, a2) = forward(p, X, din, h, dout)
    d3 = (Y - T) / B
    d2 = (d3 @ W3) * (1 - a2 ** 2)
    d1 = (d2 @ W2) * (1 - a1 ** 2)
    g = [d1.T @ X, d1.sum(0), d2.T @ a1, d2.sum(0), d3.T @ a2, d3.sum(0)]
    return np.concatenate([x.ravel() for x in g])


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** t, 1 - b2 ** t
    return p - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
