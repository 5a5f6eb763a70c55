// Graph build on one MI355X: readids text in HBM -> Node records in HBM.
//
// Pipeline (one HIP stream per handle, kernels in this order):
//   1. k_line_count / k_line_write   line terminators (Hadoop LineReader rules)
//   2. k_parse_lines                 one wave per line: tab fields, readId
//   3. k_pack_mates                  one wave per mate: validity + 2-bit pack
//   4. k_insert<W>                   one wave per mate: canonical k-mers into an
//                                    open-addressing table (count + neighbour
//                                    mask with atomics), one record per
//                                    occurrence in emission order
//   5. k_compact / radix sort        dense node array in KmerPointable order
//   6. replay (k_sel_*, radix sort, k_replay)
//                                    JDK-7 HashSet edge order for nodes whose
//                                    edge lists hold >= 2 k-mers
//   7. heads (k_heads, radix sort)   TreeSet order of 5' read heads
//   8. k_ser_size / k_ser_write      Node bytes, Java layout, big endian
// Each kernel cites the reference code whose semantics it reproduces.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/genomix_gx.h"
#include "gx_kmer.h"
#include "gx_prims.h"
#include "gx_sizer.h"
#include "../../include/genomix_batch.h"
#include "../../include/genomix_stream.h"
#include "../../include/genomix_clean.h"
#include "../../include/genomix_stats.h"

namespace gx {

struct GxError : std::runtime_error {
  int code;
  GxError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum : uint32_t { ERR_FIELDS = 1, ERR_READID = 2, ERR_READID_RANGE = 3, ERR_SHORT = 4, ERR_TABLE = 6 };

// Device error word: (line << 8) | code; the smallest line wins, as the
// reference aborts on the first bad line.
__device__ inline void report(unsigned long long* err, uint64_t line, uint32_t code) {
  atomicMin(err, (unsigned long long)((line << 8) | code));
}

// ---------------------------------------------------------------------------
// 1. line terminators: '\n', '\r' not followed by '\n', and the end of the
//    buffer when the last line has no terminator (Hadoop 0.20.2 LineReader).
// ---------------------------------------------------------------------------

constexpr int kLB = 16;                 // bytes per thread
constexpr int kLTile = kBlock * kLB;    // bytes per block

__device__ inline uint32_t count_terms(const uint8_t* __restrict__ t, size_t n, size_t i0, uint32_t* bits,
                                       unsigned long long* nonascii) {
  uint32_t c = 0, m = 0;
#pragma unroll
  for (int j = 0; j < kLB; j++) {
    size_t i = i0 + j;
    bool term = false;
    if (i < n) {
      uint8_t ch = t[i];
      if ((ch & 0x80) && nonascii) atomicMin(nonascii, (unsigned long long)i);
      term = ch == '\n' || (ch == '\r' && (i + 1 == n || t[i + 1] != '\n'));
    } else if (i == n) {
      term = n > 0 && t[n - 1] != '\n' && t[n - 1] != '\r';
    }
    if (term) { c++; m |= 1u << j; }
  }
  *bits = m;
  return c;
}

__global__ void __launch_bounds__(kBlock) k_line_count(const uint8_t* __restrict__ t, size_t n,
                                                       uint64_t* __restrict__ blockcnt,
                                                       unsigned long long* nonascii) {
  size_t i0 = (size_t)blockIdx.x * kLTile + (size_t)threadIdx.x * kLB;
  uint32_t bits;
  uint64_t c = count_terms(t, n, i0, &bits, nonascii);
  uint64_t tot;
  (void)block_exclusive_scan<uint64_t>(c, &tot);
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock) k_line_write(const uint8_t* __restrict__ t, size_t n,
                                                       const uint64_t* __restrict__ blockbase,
                                                       uint64_t* __restrict__ term,
                                                       unsigned long long* nonascii) {
  size_t i0 = (size_t)blockIdx.x * kLTile + (size_t)threadIdx.x * kLB;
  uint32_t bits;
  uint64_t c = count_terms(t, n, i0, &bits, nonascii);
  uint64_t tot;
  uint64_t pos = block_exclusive_scan<uint64_t>(c, &tot) + blockbase[blockIdx.x];
  while (bits) {
    int j = __builtin_ctz(bits);
    bits &= bits - 1;
    term[pos++] = i0 + j;
  }
}

// ---------------------------------------------------------------------------
// 0. FASTQ ingest: GenomixDriver.convertAndUploadFastqToHDFS
//    (GenomixDriver.java:665-714).  Line i (0-based, BufferedReader.readLine
//    terminators) with i % 4 == 1 becomes "<i+1>\t<mate1.trim()>\n", or with a
//    second file "<i+1>\t<mate1.trim()>\t<mate2.trim()>\n"; String.trim drops
//    leading and trailing bytes <= ' '.  The result is the library-N.readids
//    text the graph build parses.
// ---------------------------------------------------------------------------

__device__ inline void fq_line(const uint8_t* __restrict__ t, const uint64_t* __restrict__ term, uint64_t i,
                               uint64_t& s, uint64_t& e) {
  s = i ? term[i - 1] + 1 : 0;
  e = term[i];
  while (s < e && t[s] <= ' ') s++;
  while (e > s && t[e - 1] <= ' ') e--;
}

__device__ inline uint32_t ndigits(uint64_t v) {
  uint32_t d = 1;
  while (v >= 10) { v /= 10; d++; }
  return d;
}

__global__ void k_fq_size(uint64_t nrec, const uint8_t* __restrict__ t1, const uint64_t* __restrict__ term1,
                          const uint8_t* __restrict__ t2, const uint64_t* __restrict__ term2,
                          uint64_t* __restrict__ size) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < nrec; r += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t i = 4 * r + 1, s, e;
    fq_line(t1, term1, i, s, e);
    uint64_t z = ndigits(i + 1) + 1 + (e - s) + 1;
    if (t2) {
      fq_line(t2, term2, i, s, e);
      z += 1 + (e - s);
    }
    size[r] = z;
  }
}

// One wave per record; lanes copy the sequence bytes.
__global__ void __launch_bounds__(kBlock) k_fq_write(uint64_t nrec, const uint8_t* __restrict__ t1,
                                                     const uint64_t* __restrict__ term1,
                                                     const uint8_t* __restrict__ t2,
                                                     const uint64_t* __restrict__ term2,
                                                     const uint64_t* __restrict__ off, uint8_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t r = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); r < nrec; r += nw) {
    uint64_t i = 4 * r + 1, id = i + 1;
    uint8_t* p = out + off[r];
    uint32_t nd = ndigits(id);
    if (lane < (int)nd) {
      uint64_t v = id;
      for (uint32_t j = 0; j + 1 + lane < nd; j++) v /= 10;
      p[lane] = (uint8_t)('0' + v % 10);
    }
    if (lane == 0) p[nd] = '\t';
    p += nd + 1;
    for (int f = 0; f < (t2 ? 2 : 1); f++) {
      const uint8_t* t = f ? t2 : t1;
      uint64_t s, e;
      fq_line(t, f ? term2 : term1, i, s, e);
      for (uint64_t x = s + lane; x < e; x += 64) p[x - s] = t[x];
      p += e - s;
      if (lane == 0) *p = (f == 0 && t2) ? '\t' : '\n';
      p += 1;
    }
  }
}

// ---------------------------------------------------------------------------
// 2. fields of each line: ReadsKeyValueParserFactory.parse (:95-148).
//    value.toString().split("\\t") drops trailing empty strings; 2 or 3
//    fields are required, the first must parse as a Java long.
// ---------------------------------------------------------------------------

struct LineInfo {
  int64_t rid;
};

__global__ void __launch_bounds__(kBlock) k_parse_lines(const uint8_t* __restrict__ t, size_t n,
                                                        const uint64_t* __restrict__ term, uint64_t nlines,
                                                        int64_t* __restrict__ rid, uint64_t* __restrict__ mstart,
                                                        uint32_t* __restrict__ mlen, unsigned long long* err) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t line = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); line < nlines; line += nw) {
    uint64_t s = line ? term[line - 1] + 1 : 0;
    uint64_t e = term[line];
    if (e < n && t[e] == '\n' && e > s && t[e - 1] == '\r') e--;
    uint64_t ntabs = 0;
    int64_t t1 = -1, t2 = -1, lastnt = (int64_t)s - 1;
    uint32_t head = 0;  // the line's first 64 bytes, one per lane (the readId is read from here)
    // four 64-byte chunks per pass, their loads issued together
    for (uint64_t b0 = s; b0 < e; b0 += 256) {
      uint8_t chs[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint64_t pos = b0 + 64 * q + lane;
        chs[q] = pos < e ? t[pos] : 0;
      }
      if (b0 == s) head = chs[0];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint64_t b = b0 + 64 * q;
        if (b >= e) break;
        const bool in = b + lane < e;
        const uint8_t ch = chs[q];
        uint64_t mt = __ballot(in && ch == '\t');
        uint64_t mn = __ballot(in && ch != '\t');
        ntabs += __popcll(mt);
        if (t1 < 0 && mt) { t1 = (int64_t)(b + __builtin_ctzll(mt)); mt &= mt - 1; }
        if (t2 < 0 && mt) { t2 = (int64_t)(b + __builtin_ctzll(mt)); }
        if (mn) lastnt = (int64_t)(b + 63 - __builtin_clzll(mn));
      }
    }
    uint64_t e2 = (uint64_t)(lastnt + 1);           // trailing tabs stripped
    uint64_t nt = ntabs - (e - e2);                  // tabs inside [s, e2)
    uint64_t f1s = 0, f1e = 0, f2s = 0, f2e = 0;
    bool ok = e2 > s && (nt == 1 || nt == 2);
    int64_t id = 0;
    if (ok) {
      f1s = (uint64_t)t1 + 1;
      f1e = nt == 2 ? (uint64_t)t2 : e2;
      if (nt == 2) { f2s = (uint64_t)t2 + 1; f2e = e2; }
      // Long.parseLong of field 0 (lane-uniform loop; the field is short).
      uint64_t fs = s, fe = (uint64_t)t1;
      // byte i of the line: from the lanes holding the first 64, else memory
      auto byte_at = [&](uint64_t i) -> uint8_t {
        return i - s < 64 ? (uint8_t)__shfl((int)head, (int)(i - s), 64) : t[i];
      };
      bool neg = false, good = fe > fs;
      uint64_t i = fs;
      if (good) {
        const uint8_t c0 = byte_at(i);
        if (c0 == '-' || c0 == '+') { neg = c0 == '-'; i++; good = i < fe; }
      }
      unsigned __int128 v = 0;
      for (; good && i < fe; i++) {
        uint8_t ch = byte_at(i);
        if (ch < '0' || ch > '9') { good = false; break; }
        v = v * 10 + (ch - '0');
        if (v > ((unsigned __int128)1 << 63)) good = false;
      }
      if (good && !neg && v == ((unsigned __int128)1 << 63)) good = false;
      if (!good) {
        if (lane == 0) report(err, line, ERR_READID);
        ok = false;
      } else {
        id = neg ? (int64_t)(0 - (uint64_t)v) : (int64_t)(uint64_t)v;
      }
    } else if (lane == 0) {
      report(err, line, ERR_FIELDS);
    }
    if (lane == 0) {
      rid[line] = id;
      mstart[2 * line] = f1s;
      mlen[2 * line] = ok ? (uint32_t)(f1e - f1s) : 0;
      mstart[2 * line + 1] = f2s;
      mlen[2 * line + 1] = ok ? (uint32_t)(f2e - f2s) : 0;
    }
  }
}

// ---------------------------------------------------------------------------
// 3. per mate: "[ACGTacgt]+" validity (:125-128,139-141) and 2-bit packing
//    with GeneCode.getCodeFromSymbol (GeneCode.java:29-50; other bytes -> A),
//    base i at bits 2i of 64-bit word i/32, zero padded.
// ---------------------------------------------------------------------------

__device__ inline uint64_t spread32(uint32_t x32) {
  uint64_t x = x32;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

__device__ inline uint32_t code_of(uint8_t ch) {
  switch (ch) {
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
  }
  return 0;
}
__device__ inline bool is_acgt(uint8_t ch) {
  switch (ch) {
    case 'A': case 'C': case 'G': case 'T': case 'a': case 'c': case 'g': case 't': return true;
  }
  return false;
}

__global__ void __launch_bounds__(kBlock) k_pack_mates(const uint8_t* __restrict__ t, uint64_t nmates,
                                                       const uint64_t* __restrict__ mstart,
                                                       const uint32_t* __restrict__ mlen,
                                                       const uint64_t* __restrict__ wbase,
                                                       const int64_t* __restrict__ rid, int k,
                                                       uint64_t* __restrict__ words, uint8_t* __restrict__ valid,
                                                       uint64_t* __restrict__ nocc, unsigned long long* err) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t m = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); m < nmates; m += nw) {
    uint32_t L = mlen[m];
    uint64_t fs = mstart[m];
    uint64_t* w = words + wbase[m];
    bool bad = false;
    for (uint32_t b = 0; b < L; b += 64) {
      uint32_t i = b + lane;
      uint8_t ch = i < L ? t[fs + i] : 'A';
      uint32_t c = i < L ? code_of(ch) : 0;
      bad |= __ballot(i < L && !is_acgt(ch)) != 0;
      uint64_t lo = __ballot(c & 1), hi = __ballot(c >> 1);
      uint64_t w0 = spread32((uint32_t)lo) | (spread32((uint32_t)hi) << 1);
      uint64_t w1 = spread32((uint32_t)(lo >> 32)) | (spread32((uint32_t)(hi >> 32)) << 1);
      uint32_t wi = b / 32;
      if (lane == 0) w[wi] = w0;
      if (lane == 1 && b + 32 < L) w[wi + 1] = w1;
    }
    bool v = L > 0 && !bad;
    uint64_t occ = 0;
    if (v) {
      // ReadHeadInfo.makeUUID readId check (ReadHeadInfo.java:110-113, 29
      // bits) and SplitReads' length check (ReadsKeyValueParserFactory.java:152-155).
      int64_t id = rid[m >> 1];
      if (id < 0 || id >= (1ll << 29)) {
        if (lane == 0) report(err, m >> 1, ERR_READID_RANGE);
      } else if ((uint32_t)k >= L) {
        if (lane == 0) report(err, m >> 1, ERR_SHORT);
      } else {
        occ = L - k + 1;
      }
    }
    if (lane == 0) {
      valid[m] = v ? 1 : 0;
      nocc[m] = occ;
    }
  }
}

// ---------------------------------------------------------------------------
// 4. canonical k-mers into the table.  Per occurrence (SplitReads,
//    ReadsKeyValueParserFactory.java:150-196): key = min(V, R), dir F iff
//    V <= R; neighbours become an 8-bit mask relative to the canonical
//    node (bit c: successor X[1:]+c, bit 4+c: predecessor c+X[:-1]),
//    equivalent to setEdgesForCurAndNext (:209-233) over both the cur and
//    the mirrored next side.  Coverage = number of occurrences
//    (AggregateKmerAggregateFactory.java:124,144).
//    The occurrence record code byte: bits 0-1 successor letter, bit 2
//    successor present, bits 3-4 predecessor letter, bit 5 predecessor
//    present, bit 6 occurrence read in reverse (flipped).
// ---------------------------------------------------------------------------

struct Slot1 {
  uint64_t key;
  uint32_t count;
  uint32_t mask;
};
struct Slot2 {
  uint64_t hi;
  uint64_t lo;
  uint32_t count;
  uint32_t mask;
  uint64_t pad;
};
constexpr uint64_t kEmpty = ~0ull;
constexpr uint64_t kPend = 1ull << 63;

__device__ inline uint64_t ld_relaxed(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ld_relaxed32(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inserts key; returns the slot index (or ~0 when the probe limit was hit,
// i.e. the table is too full) and whether this call claimed the slot.
// Home slot by multiply-shift range reduction (any capacity, not only powers
// of two), linear probing with wrap-around.
__device__ inline uint64_t slot_home(uint64_t h, uint64_t cap) { return __umul64hi(h, cap); }
__device__ inline uint64_t slot_next(uint64_t s, uint64_t cap) { return s + 1 == cap ? 0 : s + 1; }

template <int W>
__device__ inline uint64_t table_insert(void* table, uint64_t cap, KV key, uint64_t limit, bool* claimed) {
  uint64_t s = slot_home(khash<W>(key), cap);
  *claimed = false;
  if (W == 1) {
    Slot1* T = (Slot1*)table;
    for (uint64_t probe = 0; probe < limit; probe++) {
      uint64_t cur = ld_relaxed(&T[s].key);
      if (cur == key.lo) return s;
      if (cur == kEmpty) {
        uint64_t prev = atomicCAS((unsigned long long*)&T[s].key, (unsigned long long)kEmpty,
                                  (unsigned long long)key.lo);
        if (prev == kEmpty) { *claimed = true; return s; }
        if (prev == key.lo) return s;
      }
      s = slot_next(s, cap);
    }
    return ~0ull;
  } else {
    // Two-word keys without a 128-bit CAS: claim the slot by CAS-ing the
    // high word from EMPTY to (hi | PENDING), store the low word, wait for
    // that store to complete, then publish hi.  Readers compare lo only
    // after reading a published hi equal to the key's.  Readers that see
    // PENDING retry the same slot; the claiming lane publishes within the
    // same loop iteration, so lanes of one wave never wait on each other.
    // A violated ordering could only show as one key in two slots, which
    // the node pass checks for (k_dup_check) and reports as an error.
    Slot2* T = (Slot2*)table;
    uint64_t probe = 0;
    while (probe < limit) {
      uint64_t cur = ld_relaxed(&T[s].hi);
      if (cur == kEmpty) {
        uint64_t prev = atomicCAS((unsigned long long*)&T[s].hi, (unsigned long long)kEmpty,
                                  (unsigned long long)(key.hi | kPend));
        if (prev == kEmpty) {
          __hip_atomic_store(&T[s].lo, key.lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          // lo must be visible device-wide before hi loses PENDING.  Both
          // are device-coherent (agent-scope atomic) stores, so waiting for
          // the first one's completion orders them; an agent-scope release
          // store would instead write back the whole L2 on every claim.
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          __builtin_amdgcn_s_waitcnt(0);
          __atomic_signal_fence(__ATOMIC_SEQ_CST);
          __hip_atomic_store(&T[s].hi, key.hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          *claimed = true;
          return s;
        }
        cur = prev;
      }
      if (cur & kPend) {
        probe++;  // bounded wait on a slot being published
        continue;
      }
      if (cur == key.hi) {
        // issued only after the published hi has been read (the compare
        // needs it), so it sees the claimer's completed lo store
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (ld_relaxed(&T[s].lo) == key.lo) return s;
      }
      s = slot_next(s, cap);
      probe++;
    }
    return ~0ull;
  }
}

constexpr uint32_t kEnd = 0xffffffffu;

// Read-only lookup of a key that is known to be in the table.
template <int W>
__device__ inline uint64_t table_find(const void* table, uint64_t cap, KV key) {
  uint64_t s = slot_home(khash<W>(key), cap);
  for (uint64_t probe = 0; probe < cap; probe++) {
    if (W == 1) {
      uint64_t cur = ((const Slot1*)table)[s].key;
      if (cur == key.lo) return s;
      if (cur == kEmpty) return ~0ull;
    } else {
      uint64_t cur = ((const Slot2*)table)[s].hi;
      if (cur == key.hi && ((const Slot2*)table)[s].lo == key.lo) return s;
      if (cur == kEmpty) return ~0ull;
    }
    s = slot_next(s, cap);
  }
  return ~0ull;
}

template <int W>
__device__ inline uint32_t node_of_key(const void* table, uint64_t cap, KV key) {
  uint64_t s = table_find<W>(table, cap, key);
  if (s == ~0ull) return kEnd;
  return W == 1 ? ((const Slot1*)table)[s].count : ((const Slot2*)table)[s].count;
}

template <int W>
__device__ inline KV node_key(const uint64_t* klo, const uint64_t* khi, uint64_t i) {
  KV v;
  v.lo = klo[i];
  v.hi = W == 2 ? khi[i] : 0;
  return v;
}

template <int W>
__device__ inline KV load_kmer(const uint64_t* __restrict__ w, uint32_t p, int k) {
  // bits [2p, 2p+2k) of the packed stream
  uint32_t bit = 2 * p;
  uint32_t wi = bit >> 6, sh = bit & 63;
  KV v;
  uint64_t a = w[wi];
  uint64_t b = (2 * k + sh > 64) ? w[wi + 1] : 0;
  v.lo = sh ? (a >> sh) | (b << (64 - sh)) : a;
  if (W == 2) {
    uint64_t c = (2 * k + sh > 128) ? w[wi + 2] : 0;
    v.hi = sh ? (b >> sh) | (c << (64 - sh)) : b;
  } else {
    v.hi = 0;
  }
  KV m = kmask<W>(k);
  v.lo &= m.lo;
  v.hi &= m.hi;
  return v;
}

// Edge types in EDGETYPE order (EDGETYPE.java:6-9).  The table's neighbour
// mask has bit 4t+x set when list t holds the neighbour with letter x:
// successor X[1:]+x for FF (stored as is) and FR (stored reverse-
// complemented), predecessor x+X[:-1] for RR (as is) and RF (reverse-
// complemented).  For odd k the type follows from the neighbour's
// canonicality; for even k a palindromic neighbour is stored under either
// type depending on the strand of the read, so the type is recorded per
// occurrence.  Bit 31 of the mask marks nodes whose lists need the JDK-7
// replay (set after the insert).
enum : uint32_t { ET_FF = 0, ET_FR = 1, ET_RF = 2, ET_RR = 3 };
constexpr uint32_t kNeedsBit = 1u << 31;

// Occurrence code: low nibble = successor-side entry, high nibble =
// predecessor-side entry; per nibble bit 3 present, bit 2 FR (resp. RF),
// bits 0-1 letter.
__device__ inline uint32_t nibble(uint32_t t, uint32_t x) {
  return 8u | ((t == ET_FR || t == ET_RF) ? 4u : 0u) | x;
}

// Canonical key of the k-mer at position p of a mate (L bases, packed
// words w), its occurrence code and its neighbour bits (bit 4t + x).
// The neighbours' reverse complements come from the k-mer's own:
// rc(w[1:]+c) = comp(c)+rc(w)[:-1] and rc(c+w[:-1]) = rc(w)[1:]+comp(c).
// The k-mer's words also hold its neighbour letters (the previous one in
// the word before when the k-mer starts a word), so each word is loaded
// once.  CODES = false: the key alone.
template <int W, bool CODES = true>
__device__ inline KV occ_info(const uint64_t* __restrict__ w, uint32_t L, uint32_t p, int k, uint32_t* code_out,
                              uint32_t* bits_out) {
  // bits [sh, end) of words wi, wi + 1, wi + 2 hold the k-mer; the next
  // letter is at bit end (read when p + k < L), the previous one at sh - 2
  const uint32_t wi = p >> 5, sh = 2 * (p & 31), end = sh + 2 * (uint32_t)k;
  const bool has_next = CODES && p + k < L, has_prev = CODES && p > 0;
  const uint64_t wa = w[wi];
  const uint64_t wb = (end > 64 || (has_next && end >= 64)) ? w[wi + 1] : 0;
  const uint64_t wc = (W == 2 && (end > 128 || (has_next && end >= 128))) ? w[wi + 2] : 0;
  const uint64_t wp = (has_prev && sh == 0) ? w[wi - 1] : 0;
  KV v;
  v.lo = sh ? (wa >> sh) | (wb << (64 - sh)) : wa;
  v.hi = W == 2 ? (sh ? (wb >> sh) | (wc << (64 - sh)) : wb) : 0;
  const KV km = kmask<W>(k);
  v.lo &= km.lo;
  v.hi &= km.hi;
  const KV rv = revcomp<W>(v, k);
  const bool fwd = kle<W>(v, rv);  // ties (palindromes) read forward
  KV key = fwd ? v : rv;
  uint32_t code = 0, bits = 0;
  if (!CODES) {
    *code_out = 0;
    *bits_out = 0;
    return key;
  }
  if (has_next) {  // the read's next k-mer w' = w[1:]+c
    const uint32_t c = (uint32_t)((end < 64 ? wa >> end : end < 128 ? wb >> (end - 64) : wc >> (end - 128)) & 3);
    KV nx = succ_of<W>(v, k, c);
    bool nfwd = kle<W>(nx, pred_of<W>(rv, k, 3 - c));
    // (cur F, next F) -> FF w'; (F, R) -> FR rc(w'); (R, F) -> RF w';
    // (R, R) -> RR rc(w')   (setEdgesForCurAndNext, cur side)
    uint32_t t = fwd ? (nfwd ? ET_FF : ET_FR) : (nfwd ? ET_RF : ET_RR);
    uint32_t x = fwd ? c : 3 - c;  // letter relative to the canonical node
    code |= nibble(t, x) << (fwd ? 0 : 4);
    bits |= 1u << (4 * t + x);
  }
  if (has_prev) {  // the read's previous k-mer w'' = c+w[:-1]
    const uint32_t c = (uint32_t)((sh ? wa >> (sh - 2) : wp >> 62) & 3);
    KV pv = pred_of<W>(v, k, c);
    bool pfwd = kle<W>(pv, succ_of<W>(rv, k, 3 - c));
    // (prev F, cur F) -> RR w''; (F, R) -> FR w''; (R, F) -> RF rc(w'');
    // (R, R) -> FF rc(w'')   (setEdgesForCurAndNext, next side)
    uint32_t t = fwd ? (pfwd ? ET_RR : ET_RF) : (pfwd ? ET_FR : ET_FF);
    uint32_t x = fwd ? c : 3 - c;
    code |= nibble(t, x) << (fwd ? 4 : 0);
    bits |= 1u << (4 * t + x);
  }
  *code_out = code;
  *bits_out = bits;
  return key;
}

// Neighbour bits of an occurrence code (the inverse of the nibble packing).
__device__ inline uint32_t code_bits(uint32_t code) {
  uint32_t bits = 0, sn = code & 0xF, pn = (code >> 4) & 0xF;
  if (sn & 8) bits |= 1u << (4 * ((sn & 4) ? ET_FR : ET_FF) + (sn & 3));
  if (pn & 8) bits |= 1u << (4 * ((pn & 4) ? ET_RF : ET_RR) + (pn & 3));
  return bits;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_insert(uint64_t nmates, const uint32_t* __restrict__ mlen,
                                                   const uint64_t* __restrict__ wbase,
                                                   const uint64_t* __restrict__ words,
                                                   const uint64_t* __restrict__ occbase,
                                                   const uint64_t* __restrict__ nocc, int k, void* table,
                                                   uint64_t cap, uint64_t limit,
                                                   uint32_t* __restrict__ occ_slot, uint8_t* __restrict__ occ_code,
                                                   unsigned long long* err) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t m = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); m < nmates; m += nw) {
    uint64_t nk = nocc[m];
    if (nk == 0) continue;
    uint32_t L = mlen[m];
    const uint64_t* w = words + wbase[m];
    uint64_t ob = occbase[m];
    for (uint32_t p = lane; p < nk; p += 64) {
      uint32_t code, bits;
      KV key = occ_info<W>(w, L, p, k, &code, &bits);
      bool claimed;
      uint64_t s = table_insert<W>(table, cap, key, limit, &claimed);
      (void)claimed;
      if (s == ~0ull) {
        report(err, ~0ull >> 8, ERR_TABLE);
        s = 0;
      } else {
        uint32_t* cnt = W == 1 ? &((Slot1*)table)[s].count : &((Slot2*)table)[s].count;
        uint32_t* msk = W == 1 ? &((Slot1*)table)[s].mask : &((Slot2*)table)[s].mask;
        atomicAdd(cnt, 1u);
        if ((ld_relaxed32(msk) & bits) != bits) atomicOr(msk, bits);
      }
      occ_slot[ob + p] = (uint32_t)s;
      occ_code[ob + p] = (uint8_t)code;
    }
  }
}

// HyperLogLog sketch (Flajolet et al. 2007; 2^12 registers) of the
// canonical k-mers of the first mates, for the learned table sizing.  The
// register of a k-mer is the top 12 bits of its 64-bit hash, its value the
// position of the first set bit in the remaining 52 (1..53).
constexpr int kHllBits = 12, kHllRegs = 1 << kHllBits;

template <int W>
__device__ inline uint64_t hll_hash(KV key) {
  return mix64(khash<W>(key) + 0x632be59bd9b4e019ull);
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_hll(uint64_t nmates, const uint64_t* __restrict__ wbase,
                                                const uint64_t* __restrict__ words,
                                                const uint64_t* __restrict__ nocc, int k,
                                                uint32_t* __restrict__ regs) {
  __shared__ uint32_t r[kHllRegs];
  for (int i = threadIdx.x; i < kHllRegs; i += kBlock) r[i] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t m = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); m < nmates; m += nw) {
    uint64_t nk = nocc[m];
    const uint64_t* w = words + wbase[m];
    for (uint32_t p = lane; p < nk; p += 64) {
      KV key = canon<W>(load_kmer<W>(w, p, k), k, nullptr);
      uint64_t h = hll_hash<W>(key);
      uint32_t idx = (uint32_t)(h >> (64 - kHllBits));
      uint64_t rest = (h << kHllBits) | (1ull << (kHllBits - 1));
      atomicMax(&r[idx], (uint32_t)__builtin_clzll(rest) + 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHllRegs; i += kBlock)
    if (r[i]) atomicMax(&regs[i], r[i]);
}

template <int W>
__global__ void k_table_clear(void* table, uint64_t cap) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += (uint64_t)gridDim.x * blockDim.x) {
    if (W == 1) {
      Slot1* T = (Slot1*)table;
      T[i].key = kEmpty;
      T[i].count = 0;
      T[i].mask = 0;
    } else {
      Slot2* T = (Slot2*)table;
      T[i].hi = kEmpty;
      T[i].lo = 0;
      T[i].count = 0;
      T[i].mask = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// 5. dense node array in KmerPointable order: bytes compared from the last
//    to the first (KmerPointable.java:94-107), i.e. the integer whose byte i
//    is big-endian key byte i.  The claimed-slot list gives the nodes
//    without scanning the table; the radix sort fixes their order.
// ---------------------------------------------------------------------------

template <int W>
__device__ inline KV slot_key(const void* table, uint64_t i) {
  KV v;
  if (W == 1) {
    v.lo = ((const Slot1*)table)[i].key;
    v.hi = 0;
  } else {
    v.lo = ((const Slot2*)table)[i].lo;
    v.hi = ((const Slot2*)table)[i].hi;
  }
  return v;
}

__device__ inline uint64_t sortkey64(uint64_t v, int nb) { return __builtin_bswap64(v) >> (64 - 8 * nb); }
__device__ inline Key128 sortkey128(KV v, int nb) {
  // bswap over 128 bits, then shift right by 128 - 8nb
  uint64_t lo = __builtin_bswap64(v.hi), hi = __builtin_bswap64(v.lo);
  int sh = 128 - 8 * nb;
  Key128 r;
  if (sh >= 64) {
    r.lo = hi >> (sh - 64);
    r.hi = 0;
  } else if (sh == 0) {
    r.lo = lo;
    r.hi = hi;
  } else {
    r.lo = (lo >> sh) | (hi << (64 - sh));
    r.hi = hi >> sh;
  }
  return r;
}

template <int W>
__device__ inline bool slot_used(const void* table, uint64_t i) {
  return W == 1 ? ((const Slot1*)table)[i].key != kEmpty : ((const Slot2*)table)[i].hi != kEmpty;
}

constexpr int kCItems = 16;
constexpr int kCTile = kBlock * kCItems;

// Occupied slots, striped over the block so every load is coalesced; the
// order of the list is irrelevant (the radix sort fixes the node order).
template <int W>
__global__ void __launch_bounds__(kBlock) k_compact_count(const void* table, uint64_t cap, uint64_t* blockcnt) {
  uint64_t c = 0;
  uint64_t base = (uint64_t)blockIdx.x * kCTile;
  for (int i = 0; i < kCItems; i++) {
    uint64_t j = base + (uint64_t)i * kBlock + threadIdx.x;
    if (j < cap && slot_used<W>(table, j)) c++;
  }
  uint64_t tot;
  (void)block_exclusive_scan<uint64_t>(c, &tot);
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = tot;
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_compact_write(const void* table, uint64_t cap,
                                                          const uint64_t* __restrict__ blockbase,
                                                          uint32_t* __restrict__ claimed) {
  uint64_t base = (uint64_t)blockIdx.x * kCTile;
  uint64_t c = 0;
  for (int i = 0; i < kCItems; i++) {
    uint64_t j = base + (uint64_t)i * kBlock + threadIdx.x;
    if (j < cap && slot_used<W>(table, j)) c++;
  }
  uint64_t tot;
  uint64_t pos = block_exclusive_scan<uint64_t>(c, &tot) + blockbase[blockIdx.x];
  for (int i = 0; i < kCItems; i++) {
    uint64_t j = base + (uint64_t)i * kBlock + threadIdx.x;
    if (j < cap && slot_used<W>(table, j)) claimed[pos++] = (uint32_t)j;
  }
}

template <int W, class SK>
__global__ void k_compact_list(const void* table, const uint32_t* __restrict__ claimed, uint64_t nn, int nb,
                               SK* __restrict__ skey, uint32_t* __restrict__ sval) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    KV v = slot_key<W>(table, claimed[i]);
    if constexpr (W == 1) skey[i] = sortkey64(v.lo, nb);
    else skey[i] = sortkey128(v, nb);
    sval[i] = (uint32_t)i;
  }
}

// Node arrays in output order.  The slot's count field is replaced by the
// node's output rank once its count has been copied out.
// ---------------------------------------------------------------------------
// 4'-5'. Sort-partitioned aggregation (the default method).  Every k-mer
//    occurrence becomes a record (KmerPointable sort key, occurrence index
//    << 8 | code); the records are sorted by (key, occurrence): two stable
//    8-bit radix passes on the top 16 key bits -- the first fused with the
//    k-mer extraction -- then a bitonic sort of each 16-bit bucket in LDS
//    (buckets above kBucketCap records go through a global radix sort).
//    Nodes are the runs of equal keys, already in output order; a run's
//    codes are its occurrences' records in emission order, as the JDK-7
//    replay needs them.  All HBM traffic is streaming; there is no
//    random-access table in the aggregation.
// ---------------------------------------------------------------------------

template <int W> struct SortKeyOf { using T = uint64_t; };
template <> struct SortKeyOf<2> { using T = Key128; };

__device__ inline void make_sort_key(KV v, int nb, uint64_t& out) { out = sortkey64(v.lo, nb); }
__device__ inline void make_sort_key(KV v, int nb, Key128& out) { out = sortkey128(v, nb); }

// inverse of make_sort_key
__device__ inline KV key_of_sort(uint64_t sk, int nb) {
  KV v;
  v.lo = __builtin_bswap64(sk << (64 - 8 * nb));
  v.hi = 0;
  return v;
}
__device__ inline KV key_of_sort(const Key128& r, int nb) {
  int sh = 128 - 8 * nb;  // 0..64 for two-word keys
  uint64_t lo, hi;
  if (sh == 0) { lo = r.lo; hi = r.hi; }
  else if (sh >= 64) { hi = r.lo << (sh - 64); lo = 0; }
  else { hi = (r.hi << sh) | (r.lo >> (64 - sh)); lo = r.lo << sh; }
  KV v;
  v.lo = __builtin_bswap64(hi);
  v.hi = __builtin_bswap64(lo);
  return v;
}

__device__ inline bool sk_eq(uint64_t a, uint64_t b) { return a == b; }
__device__ inline bool sk_eq(const Key128& a, const Key128& b) { return a.lo == b.lo && a.hi == b.hi; }

// Record-sort bucket: the top `bkb` bytes of the sort key (2, or 3 for
// passes of more than 2^28 records, so buckets stay LDS-sized; 1 when the
// key has one byte).
template <class SK>
__device__ inline uint32_t sort_bucket(const SK& sk, int nb, int bkb) {
  uint32_t v = 0;
  for (int i = 1; i <= bkb; i++) v = (v << 8) | key_byte(sk, nb - i);
  return v;
}

// Key -> node index, built from the sorted nodes without global atomics:
// the nodes of each key bucket (the top bkb bytes of the sort key: 2, or 3
// above 2^26 nodes) are contiguous (dir[b] = first node), and each bucket
// owns a private open-addressing region of two slots per node, filled in
// LDS by one block.  A lookup is the bucket's two dir entries plus one or
// two slot probes.
struct NodeIndex {
  const uint32_t* dir;  // nbk + 1 entries
  const void* slots;    // Slot1 / Slot2, 2 per node; count = node id
  int nb;
  uint32_t nbk;
  int bkb;
};

template <int W>
__device__ inline uint32_t node_bucket(KV key, int nb, int bkb) {
  typename SortKeyOf<W>::T sk;
  make_sort_key(key, nb, sk);
  return sort_bucket(sk, nb, bkb);
}

template <int W>
__device__ inline uint32_t node_lookup(const NodeIndex& ix, KV key) {
  uint32_t b = node_bucket<W>(key, ix.nb, ix.bkb);
  uint32_t lo = ix.dir[b], hi = ix.dir[b + 1];
  if (lo >= hi) return kEnd;
  const uint64_t R = 2ull * (hi - lo), base = 2ull * lo;
  uint64_t s = slot_home(khash<W>(key), R);
  for (uint64_t probe = 0; probe < R; probe++) {
    if (W == 1) {
      const Slot1& e = ((const Slot1*)ix.slots)[base + s];
      if (e.key == key.lo) return e.count;
      if (e.key == kEmpty) return kEnd;
    } else {
      const Slot2& e = ((const Slot2*)ix.slots)[base + s];
      if (e.hi == key.hi && e.lo == key.lo) return e.count;
      if (e.hi == kEmpty) return kEnd;
    }
    s = s + 1 == R ? 0 : s + 1;
  }
  return kEnd;
}

// dir over buckets [b0, b1] (every node's bucket lies in [b0, b1)); the
// entries outside are filled by k_fill_u32.
template <int W>
__global__ void k_nix_dir(uint64_t nn, const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi, int nb,
                          int bkb, uint32_t b0, uint32_t b1, uint32_t* __restrict__ dir) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= nn; j += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = j == 0 ? b0 : node_bucket<W>(node_key<W>(klo, khi, j - 1), nb, bkb) + 1;
    uint32_t hi = j == nn ? b1 : node_bucket<W>(node_key<W>(klo, khi, j), nb, bkb);
    for (uint32_t b = lo; b <= hi && b <= b1; b++) dir[b] = (uint32_t)j;
  }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint64_t from, uint64_t to, uint32_t v) {
  for (uint64_t i = from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < to; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

constexpr int kNixLds = 4096;  // slots of one bucket region built in LDS (two-byte buckets)
constexpr int kNixLdsSmall = 1024;  // three-byte buckets: smaller regions, more blocks per CU
template <int W, int LCAP>
__global__ void __launch_bounds__(kBlock) k_nix_build(const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi,
                                                      const uint32_t* __restrict__ msk, const uint32_t* __restrict__ dir,
                                                      void* slots, uint32_t b0, uint32_t b1) {
  using S = typename std::conditional<W == 1, Slot1, Slot2>::type;
  __shared__ S lt[LCAP];
  for (uint32_t bk = b0 + blockIdx.x; bk < b1; bk += gridDim.x) {
    const uint32_t lo = dir[bk], hi = dir[bk + 1];
    if (lo >= hi) continue;
    const uint64_t R = 2ull * (hi - lo), base = 2ull * lo;
    const bool local = R <= LCAP;
    S* T = local ? lt : (S*)slots + base;
    for (uint64_t i = threadIdx.x; i < R; i += kBlock) {
      S e{};
      if (W == 1) ((Slot1*)&e)->key = kEmpty;
      else ((Slot2*)&e)->hi = kEmpty;
      T[i] = e;
    }
    __syncthreads();
    for (uint64_t n = lo + threadIdx.x; n < hi; n += kBlock) {
      KV key = node_key<W>(klo, khi, n);
      uint64_t s = slot_home(khash<W>(key), R);
      for (uint64_t probe = 0; probe < R; probe++) {
        if (W == 1) {
          Slot1* e = (Slot1*)&T[s];
          if (atomicCAS((unsigned long long*)&e->key, (unsigned long long)kEmpty, (unsigned long long)key.lo) ==
              kEmpty) {
            e->count = (uint32_t)n;
            e->mask = msk[n] & 0xFFFFu;
            break;
          }
        } else {  // distinct keys: a taken slot always holds another key
          Slot2* e = (Slot2*)&T[s];
          if (atomicCAS((unsigned long long*)&e->hi, (unsigned long long)kEmpty, (unsigned long long)key.hi) ==
              kEmpty) {
            e->lo = key.lo;
            e->count = (uint32_t)n;
            e->mask = msk[n] & 0xFFFFu;
            break;
          }
        }
        s = s + 1 == R ? 0 : s + 1;
      }
    }
    __syncthreads();
    if (local)
      for (uint64_t i = threadIdx.x; i < R; i += kBlock) ((S*)slots)[base + i] = lt[i];
    __syncthreads();  // lt is reused by the next bucket
  }
}

constexpr int kBucketCap = 8192;  // records sorted per bucket in one block (72 KB / 144 KB of LDS)
// HBM of a key-range pass, per occurrence (sizes the ranges): the records
// (two key and two neighbour-code arrays, the sorted codes) and, per node (dfrac
// nodes per occurrence), every node array of the pass -- key words, count,
// mask, first record, two index slots, replay flags, counts, offsets,
// lists and order, head counts and starts, record sizes and offsets -- and
// the record itself, estimated as three VKmers and headers.  Counted as
// live together: freeing dead buffers on demand works, but re-allocating
// 100 GB-scale buffers every pass costs more than an extra pass.
__host__ inline double pass_bytes_per_occurrence(int W, int nb, double dfrac) {
  const double rec = 2.0 * (W == 1 ? 8 : 16) + 2 + 1;
  const double node = 8.0 * W + 4 + 4 + 8 + 2.0 * (W == 1 ? 16 : 24) + 4 + 8 + 8 + 8 + 8 + 16 + 8 + 16 + 16;
  const double out = 3.0 * (4 + nb) + 16;
  return rec + dfrac * (node + out);
}
// HBM one node of a key range holds until the build ends when the records
// stay in HBM: its record (estimated as above), its node graph (key words,
// count, mask, two head counts, about one head), and its share of what is
// built over the whole graph after the last range (two node-index slots,
// head total and start), which stays allocated beside the next build's
// per-range buffers.
__host__ inline double kept_bytes_per_node(int W, int nb) {
  return 3.0 * (4 + nb) + 16 + 8.0 * W + 4 * 4 + 4 + 2.0 * (W == 1 ? 16 : 24) + 16;
}
constexpr int kIdxBits = 13;      // log2(kBucketCap)

// The bucket bytes are sorted least significant first (LSD): the fused
// extraction pass scatters by byte nb-bkb, radix passes after it by bytes
// nb-bkb+1 .. nb-1; a one-byte key has a single pass on byte 0.
constexpr int kTileMateCap = 512;

// Mate holding occurrence j: the largest m in [lo, hi] with occbase[m] <= j.
__device__ inline uint64_t mate_search(const uint64_t* __restrict__ ob, uint64_t lo, uint64_t hi, uint64_t j) {
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo + 1) / 2;
    if (ob[mid] <= j) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The mates of a tile's occurrences [j0, j1), in LDS: the tile's first and
// last mate (mr), and, when the mates fit (always, for mates of at least
// kSortTile / kTileMateCap occurrences on average), their first occurrences
// relative to the first mate's (lob) and a table giving each occurrence's
// mate (mtab): every non-empty mate marks the slot of its first occurrence
// and a running maximum over the tile fills the rest; otherwise each
// occurrence searches occbase.
struct TileMates {
  uint64_t* mr;    // [2] first and last mate; [2] = the first mate's first occurrence
  uint32_t* lob;   // [kTileMateCap] occbase[m0 + i] - occbase[m0]
  uint16_t* mtab;  // [kSortTile] mate of occurrence j0 + q, relative to m0
};

__device__ inline void tile_mates_init(const TileMates& t, uint64_t j0, uint64_t j1, uint64_t nmates,
                                       const uint64_t* __restrict__ occbase) {
  __shared__ uint16_t wmax[kBlock / 64];
  if (threadIdx.x == 0) {
    t.mr[0] = mate_search(occbase, 0, nmates - 1, j0);
    t.mr[1] = mate_search(occbase, t.mr[0], nmates - 1, j1 - 1);
    t.mr[2] = occbase[t.mr[0]];
  }
  __syncthreads();
  const uint64_t m0 = t.mr[0], m1 = t.mr[1], b0 = t.mr[2];
  if (m1 - m0 + 2 <= kTileMateCap) {
    for (uint64_t i = threadIdx.x; i <= m1 - m0 + 1; i += kBlock)
      t.lob[i] = i + m0 <= nmates ? (uint32_t)(occbase[m0 + i] - b0) : 0xffffffffu;
    for (uint32_t q = threadIdx.x; q < kSortTile; q += kBlock) t.mtab[q] = 0;
    __syncthreads();
    for (uint64_t i = threadIdx.x + 1; i <= m1 - m0; i += kBlock)
      if (t.lob[i + 1] > t.lob[i] && b0 + t.lob[i] < j1) t.mtab[b0 + t.lob[i] - j0] = (uint16_t)i;
    __syncthreads();
    // running maximum: each thread over kSortTile / kBlock consecutive slots
    constexpr int kPer = kSortTile / kBlock;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint16_t* my = t.mtab + threadIdx.x * kPer;
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < kPer; q++) run = max(run, (uint32_t)my[q]);
    uint32_t incl = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(incl, o);
      if (lane >= o) incl = max(incl, y);
    }
    if (lane == 63) wmax[wid] = (uint16_t)incl;
    __syncthreads();
    uint32_t before = __shfl_up(incl, 1);
    if (lane == 0) before = 0;
    for (int v = 0; v < wid; v++) before = max(before, (uint32_t)wmax[v]);
#pragma unroll
    for (int q = 0; q < kPer; q++) {
      before = max(before, (uint32_t)my[q]);
      my[q] = (uint16_t)before;
    }
  } else {
    __syncthreads();
  }
  __syncthreads();
}

// Canonical key (and, with CODES, neighbour code) of occurrence j of the tile.
template <int W, bool CODES>
__device__ inline KV tile_occurrence(const TileMates& t, uint64_t j0, uint64_t j, const uint64_t* __restrict__ occbase,
                                     const uint32_t* __restrict__ mlen, const uint64_t* __restrict__ wbase,
                                     const uint64_t* __restrict__ words, int k, uint32_t* code) {
  const uint64_t m0 = t.mr[0], m1 = t.mr[1];
  const bool local = m1 - m0 + 2 <= kTileMateCap;
  const uint64_t m = local ? m0 + t.mtab[j - j0] : mate_search(occbase, m0, m1, j);
  const uint64_t mb = local ? t.mr[2] + t.lob[m - m0] : occbase[m];
  uint32_t bits;
  return occ_info<W, CODES>(words + wbase[m], mlen[m], (uint32_t)(j - mb), k, code, &bits);
}

// Occurrences [j0, j1) of a tile: calls f(round, valid, j, key, code) for
// each, in rounds of kBlock consecutive occurrences.
template <int W, bool CODES, class F>
__device__ inline void for_tile_occurrences(const TileMates& t, uint64_t j0, uint64_t j1, uint64_t nmates,
                                            const uint64_t* __restrict__ occbase, const uint32_t* __restrict__ mlen,
                                            const uint64_t* __restrict__ wbase, const uint64_t* __restrict__ words,
                                            int k, F&& f) {
  tile_mates_init(t, j0, j1, nmates, occbase);
  for (int r = 0; r < kSortItems; r++) {
    uint64_t j = j0 + (uint64_t)r * kBlock + threadIdx.x;
    bool valid = j < j1;
    KV key{0, 0};
    uint32_t code = 0;
    if (valid) key = tile_occurrence<W, CODES>(t, j0, j, occbase, mlen, wbase, words, k, &code);
    f(r, valid, j, key, code);
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_occ_hist(uint64_t N, uint64_t nmates, const uint64_t* __restrict__ occbase,
                                                     const uint32_t* __restrict__ mlen,
                                                     const uint64_t* __restrict__ wbase,
                                                     const uint64_t* __restrict__ words, int k, int nb,
                                                     uint32_t* __restrict__ hist, unsigned ntiles, uint32_t tlo,
                                                     uint32_t thi, int dbyte) {
  using SK = typename SortKeyOf<W>::T;
  __shared__ uint32_t h[256];
  __shared__ uint64_t mr[3];
  __shared__ uint32_t lob[kTileMateCap];
  __shared__ uint16_t mtab[kSortTile];
  h[threadIdx.x] = 0;
  uint64_t j0 = (uint64_t)blockIdx.x * kSortTile, j1 = j0 + kSortTile < N ? j0 + kSortTile : N;
  for_tile_occurrences<W, false>(TileMates{mr, lob, mtab}, j0, j1, nmates, occbase, mlen, wbase, words, k,
                                 [&](int, bool valid, uint64_t, KV key, uint32_t) {
                                   if (valid) {
                                     SK sk;
                                     make_sort_key(key, nb, sk);
                                     const uint32_t t = key_byte(sk, nb - 1);
                                     if (t >= tlo && t < thi) atomicAdd(&h[key_byte(sk, dbyte)], 1u);
                                   }
                                 });
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of the tile's records by the lower bucket byte (same
// ranking as k_radix_scatter: wave match + lower waves' counts), staged in
// LDS in digit order and written out as one contiguous run per digit.  The
// stage holds 4 bytes per record (occurrence in the tile, neighbour code);
// the write-out extracts each record's key again from the packed mates
// (which the tile has just read, so they come from the caches), which
// halves the block's LDS and doubles the blocks a CU holds.  A record is
// its sort key and its neighbour code (one byte): the sorts after this are
// stable or break ties by the record's index in its bucket, so occurrence
// order among equal keys needs no stored occurrence number.
template <int W>
__global__ void __launch_bounds__(kBlock) k_occ_scatter(uint64_t N, uint64_t nmates,
                                                        const uint64_t* __restrict__ occbase,
                                                        const uint32_t* __restrict__ mlen,
                                                        const uint64_t* __restrict__ wbase,
                                                        const uint64_t* __restrict__ words, int k, int nb,
                                                        const uint64_t* __restrict__ offs, unsigned ntiles,
                                                        typename SortKeyOf<W>::T* __restrict__ okeys,
                                                        uint8_t* __restrict__ ocodes, uint32_t tlo, uint32_t thi,
                                                        int dbyte, uint32_t nlo, uint32_t nhi, int ndbyte,
                                                        uint32_t* __restrict__ nhist) {
  using SK = typename SortKeyOf<W>::T;
  __shared__ uint64_t run[256];
  __shared__ uint16_t lbase[256], lrun[256];  // (LDS kept under 20 KB: 8 blocks per CU)
  __shared__ uint32_t wsum[kBlock / 64];
  __shared__ uint16_t wcnt[kBlock / 64][256];
  __shared__ uint64_t mr[3];
  __shared__ uint32_t lob[kTileMateCap];
  __shared__ uint16_t mtab[kSortTile];
  __shared__ uint32_t h2[128];  // the next key range's digit counts (as k_occ_hist would give), two u16 per word
  __shared__ uint32_t sv[kSortTile];  // (occurrence - j0) << 8 | code
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  run[threadIdx.x] = offs[(size_t)threadIdx.x * ntiles + blockIdx.x];
  if (threadIdx.x < 128) h2[threadIdx.x] = 0;
  tile_digit_bases(offs, ntiles, lbase, lrun, wsum);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t j0 = (uint64_t)blockIdx.x * kSortTile, j1 = j0 + kSortTile < N ? j0 + kSortTile : N;
  const TileMates tm{mr, lob, mtab};
  for_tile_occurrences<W, true>(tm, j0, j1, nmates, occbase, mlen, wbase, words, k,
                                [&](int, bool valid, uint64_t j, KV key, uint32_t code) {
    for (int w = 0; w < kBlock / 64; w++) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    SK skey{};
    uint32_t d = 0;
    if (valid) {
      make_sort_key(key, nb, skey);
      d = key_byte(skey, dbyte);
      const uint32_t t = key_byte(skey, nb - 1);
      if (nhist && t >= nlo && t < nhi) {
        const uint32_t nd = key_byte(skey, ndbyte);
        atomicAdd(&h2[nd >> 1], 1u << (16 * (nd & 1)));  // at most kSortTile per digit
      }
      valid = t >= tlo && t < thi;  // key-range pass: other records wait for their range
    }
    uint64_t act = __ballot(valid);
    uint64_t match = act;
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t m = __ballot((d >> b) & 1);
      match &= ((d >> b) & 1) ? m : ~m;
    }
    uint32_t rank = __popcll(match & lt);
    if (valid && rank == 0) wcnt[wid][d] = (uint16_t)__popcll(match);
    __syncthreads();
    if (valid) {
      uint32_t pos = lrun[d] + rank;
      for (int w = 0; w < wid; w++) pos += wcnt[w][d];
      sv[pos] = ((uint32_t)(j - j0) << 8) | code;
    }
    __syncthreads();
    uint32_t tot = 0;
    for (int w = 0; w < kBlock / 64; w++) tot += wcnt[w][threadIdx.x];
    lrun[threadIdx.x] += tot;
  });
  __syncthreads();
  const uint32_t tn = lrun[255];  // records of this range in the tile
  for (uint32_t i = threadIdx.x; i < tn; i += kBlock) {
    const uint32_t v = sv[i];
    const uint64_t j = j0 + (v >> 8);
    uint32_t code_unused;
    SK skey;
    make_sort_key(tile_occurrence<W, false>(tm, j0, j, occbase, mlen, wbase, words, k, &code_unused), nb, skey);
    const uint32_t d = key_byte(skey, dbyte);
    const uint64_t pos = run[d] + (i - lbase[d]);
    okeys[pos] = skey;
    ocodes[pos] = (uint8_t)v;
  }
  if (nhist)
    nhist[(size_t)threadIdx.x * ntiles + blockIdx.x] = (h2[threadIdx.x >> 1] >> (16 * (threadIdx.x & 1))) & 0xffffu;
}

// Occurrences per value of the sort key's top byte (the key ranges of a
// multi-pass build are cut on it).
template <int W>
__global__ void __launch_bounds__(kBlock) k_top_hist(uint64_t N, uint64_t nmates, const uint64_t* __restrict__ occbase,
                                                     const uint32_t* __restrict__ mlen,
                                                     const uint64_t* __restrict__ wbase,
                                                     const uint64_t* __restrict__ words, int k, int nb,
                                                     unsigned long long* __restrict__ out) {
  using SK = typename SortKeyOf<W>::T;
  __shared__ uint32_t h[256];
  __shared__ uint64_t mr[3];
  __shared__ uint32_t lob[kTileMateCap];
  __shared__ uint16_t mtab[kSortTile];
  h[threadIdx.x] = 0;
  uint64_t j0 = (uint64_t)blockIdx.x * kSortTile, j1 = j0 + kSortTile < N ? j0 + kSortTile : N;
  for_tile_occurrences<W, false>(TileMates{mr, lob, mtab}, j0, j1, nmates, occbase, mlen, wbase, words, k,
                                 [&](int, bool valid, uint64_t, KV key, uint32_t) {
                                   if (valid) {
                                     SK sk;
                                     make_sort_key(key, nb, sk);
                                     atomicAdd(&h[key_byte(sk, nb - 1)], 1u);
                                   }
                                 });
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&out[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// First record of every bucket of [b0, b1) (start[b1] = N; empty buckets
// start where the next one does).  Every record's bucket lies in [b0, b1);
// entries outside [b0, b1] are not written.
template <class SK>
__global__ void k_bucket_starts(uint64_t N, const SK* __restrict__ keys, int nb, int bkb, uint32_t b0, uint32_t b1,
                                uint64_t* __restrict__ start) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= N; j += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = j == 0 ? b0 : sort_bucket(keys[j - 1], nb, bkb) + 1;
    uint32_t hi = j == N ? b1 : sort_bucket(keys[j], nb, bkb);
    for (uint32_t b = lo; b <= hi && b <= b1; b++) start[b] = j;
  }
}

// In-bucket sort element: the key bits below the bucket, then the record's
// index in the bucket (so ties keep the occurrence order).
__device__ inline uint64_t bucket_elem(uint64_t sk, int nb, int bkb, uint32_t idx) {
  int bits = 8 * (nb - bkb);  // <= 48
  uint64_t low = bits >= 64 ? sk : (sk & ((1ull << bits) - 1));
  return (low << kIdxBits) | idx;
}
__device__ inline Key128 bucket_elem(const Key128& sk, int nb, int bkb, uint32_t idx) {
  int bits = 8 * (nb - bkb);  // 40..112
  Key128 low = sk;
  if (bits < 64) { low.lo &= (1ull << bits) - 1; low.hi = 0; }
  else if (bits < 128) low.hi &= (bits == 64) ? 0ull : ((1ull << (bits - 64)) - 1);
  Key128 r;
  r.hi = (low.hi << kIdxBits) | (low.lo >> (64 - kIdxBits));
  r.lo = (low.lo << kIdxBits) | idx;
  return r;
}
__device__ inline uint32_t elem_idx(uint64_t e) { return (uint32_t)(e & (kBucketCap - 1)); }
__device__ inline uint32_t elem_idx(const Key128& e) { return (uint32_t)(e.lo & (kBucketCap - 1)); }
__device__ inline bool elem_gt(uint64_t a, uint64_t b) { return a > b; }
__device__ inline bool elem_gt(const Key128& a, const Key128& b) { return a.hi > b.hi || (a.hi == b.hi && a.lo > b.lo); }
__device__ inline void elem_max(uint64_t& e) { e = ~0ull; }
__device__ inline void elem_max(Key128& e) { e.lo = e.hi = ~0ull; }

// Bitonic exchange partner in another lane of the wave (lane ^ m).
__device__ inline uint64_t lane_xor(uint64_t x, int m) {
  const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)x, m);
  const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(x >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}
__device__ inline Key128 lane_xor(const Key128& x, int m) {
  Key128 r;
  r.lo = lane_xor(x.lo, m);
  r.hi = lane_xor(x.hi, m);
  return r;
}
// Stages whose partner is in the same thread (stride S < elements per
// thread): compile-time indices keep the elements in registers.
// The direction of element i0 + e is bit `size` of i0 when size >= EPT
// (up_t, the same for all of the thread's elements), else bit `size` of e.
template <int S, int EPT, class E>
__device__ inline void bitonic_in_thread(E* x, uint32_t size, bool up_t) {
#pragma unroll
  for (int e = 0; e < EPT; e++) {
    if (e & S) continue;
    const bool up = size >= (uint32_t)EPT ? up_t : (((uint32_t)e & size) == 0);
    if (elem_gt(x[e], x[e | S]) == up) {
      const E y = x[e];
      x[e] = x[e | S];
      x[e | S] = y;
    }
  }
}

// One block per bucket: bitonic sort of (key below the bucket, index),
// then the records are written in that order: the key and its neighbour
// code (one byte per record: the replay's code array).  Each thread holds
// EPT = CAP / NT consecutive elements in registers: stages of stride < EPT
// run inside the thread, strides < 64 * EPT across the wave's lanes
// (shuffles, no barrier), only the larger ones through LDS.  Element i
// lives at LDS slot i + i / EPT, so the threads' blocks of EPT fall on
// different banks.  Four tiers: one wave per bucket (CAP =
// kSmallBucket) over every bucket of the key range; it lists the buckets
// above that for the CAP = kBucketCap / 8 (256 threads), kBucketCap / 2
// (512) and kBucketCap (1024) blocks, which run over their lists only
// (buckets marked 4, 3, 2).  Oversized buckets (above kBucketCap) are
// marked 1 and left to the global pass.
constexpr int kSmallBucket = 256;
template <int W, int CAP, int NT>
__device__ inline void sort_one_bucket(uint32_t bk, typename SortKeyOf<W>::T* s,
                                       const typename SortKeyOf<W>::T* __restrict__ ik,
                                       const uint8_t* __restrict__ ic, typename SortKeyOf<W>::T* __restrict__ ok,
                                       uint8_t* __restrict__ oc, const uint64_t* __restrict__ bstart, uint64_t N,
                                       int nb, int bkb, uint32_t* __restrict__ big, uint32_t* __restrict__ lists,
                                       uint32_t* __restrict__ nlist, uint32_t NBK, unsigned long long* err) {
  using SK = typename SortKeyOf<W>::T;
  using E = SK;  // uint64_t or Key128
  const bool first = CAP == kSmallBucket;
  const uint64_t a = bstart[bk], e = bstart[bk + 1];
  if (a > e || e > N) {  // bucket bounds must be ordered: report, touch nothing
    if (threadIdx.x == 0) atomicMin(err, (unsigned long long)bk);
    return;
  }
  const uint32_t n = (uint32_t)(e - a);
  if (n == 0) return;
  if (first && n > CAP) {
    if (n <= kBucketCap) {  // to the tier that fits it
      if (threadIdx.x == 0) {
        const int t = n <= kBucketCap / 8 ? 0 : n <= kBucketCap / 2 ? 1 : 2;
        big[bk] = 4 - t;
        lists[(size_t)t * NBK + atomicAdd(&nlist[t], 1u)] = bk;
      }
      return;
    }
    if (threadIdx.x == 0) big[bk] = 1;
  }
  if (n > (uint32_t)CAP) {  // the global pass sorts and writes it; in a listed tier, a bad list
    if (!first && threadIdx.x == 0) atomicMin(err, (unsigned long long)bk);
    return;
  }
  if (n == 1 || nb < 2) {
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      ok[a + i] = ik[a + i];
      oc[a + i] = ic[a + i];
    }
    return;
  }
  constexpr uint32_t EPT = CAP / NT, WSPAN = 64 * EPT;  // elements per thread, per wave
  static_assert(EPT == 4 || EPT == 8, "four or eight elements per thread");
  uint32_t P = EPT;  // network width: a power of two >= n, at least one thread's share
  while (P < n) P <<= 1;
  for (uint32_t i = threadIdx.x; i < P; i += NT) {
    E x;
    if (i < n) x = bucket_elem(ik[a + i], nb, bkb, i);
    else elem_max(x);
    s[i + i / EPT] = x;
  }
  __syncthreads();
  const uint32_t t = threadIdx.x, i0 = t * EPT, slot0 = i0 + t;  // slot of element i0 + e: slot0 + e
  const bool act = i0 < P;  // threads past the network's width only take part in the barriers
  E x[EPT];
  if (act) {
#pragma unroll
    for (uint32_t e = 0; e < EPT; e++) x[e] = s[slot0 + e];
  }
  for (uint32_t size = 2; size <= P; size <<= 1) {
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      // stride >= EPT: the stride and size bits of i0 + e are those of i0,
      // so the direction is the thread's and the partner is element e of
      // thread t ^ (stride / EPT)
      const bool keep_min = ((i0 & stride) == 0) == ((i0 & size) == 0);
      if (stride >= WSPAN) {  // partner in another wave
        __syncthreads();
        if (act) {
#pragma unroll
          for (uint32_t e = 0; e < EPT; e++) s[slot0 + e] = x[e];
        }
        __syncthreads();
        if (act) {
          const uint32_t pslot0 = (t ^ (stride / EPT)) * (EPT + 1);
#pragma unroll
          for (uint32_t e = 0; e < EPT; e++) {
            const E y = s[pslot0 + e];
            if (elem_gt(x[e], y) == keep_min) x[e] = y;
          }
        }
      } else if (stride >= EPT) {  // partner in another lane (active: j < P)
        if (act) {
          const int m = (int)(stride / EPT);
#pragma unroll
          for (uint32_t e = 0; e < EPT; e++) {
            const E y = lane_xor(x[e], m);
            if (elem_gt(x[e], y) == keep_min) x[e] = y;
          }
        }
      } else if (act) {
        const bool up_t = (i0 & size) == 0;
        if (stride == 1) bitonic_in_thread<1, EPT>(x, size, up_t);
        else if (stride == 2) bitonic_in_thread<2, EPT>(x, size, up_t);
        else if constexpr (EPT > 4) bitonic_in_thread<4, EPT>(x, size, up_t);
      }
    }
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (uint32_t e = 0; e < EPT; e++) s[slot0 + e] = x[e];
  }
  __syncthreads();
  if constexpr (W == 1) {
    // the element holds the key bits below the bucket, and the bucket id
    // is the bytes above them: the key needs no second read
    const int bits = 8 * (nb - bkb);
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const E el = s[i + i / EPT];
      ok[a + i] = ((uint64_t)bk << bits) | (el >> kIdxBits);
      oc[a + i] = ic[a + elem_idx(el)];
    }
  } else {
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
      const uint32_t src = elem_idx(s[i + i / EPT]);
      ok[a + i] = ik[a + src];
      oc[a + i] = ic[a + src];
    }
  }
}

// Blocks loop over the buckets (grid x block stays below 2^32 work-items):
// the first tier over buckets [bucket0, bucket1), the others over a list.
template <int W, int CAP, int NT>
__global__ void __launch_bounds__(NT) k_bucket_sort(const typename SortKeyOf<W>::T* __restrict__ ik,
                                                        const uint8_t* __restrict__ ic,
                                                        typename SortKeyOf<W>::T* __restrict__ ok,
                                                        uint8_t* __restrict__ oc, const uint64_t* __restrict__ bstart,
                                                        uint64_t N, int nb, int bkb, uint32_t bucket0,
                                                        uint32_t count, const uint32_t* __restrict__ list,
                                                        uint32_t* __restrict__ big, uint32_t* __restrict__ lists,
                                                        uint32_t* __restrict__ nlist, uint32_t NBK,
                                                        unsigned long long* err) {
  __shared__ typename SortKeyOf<W>::T s[CAP + NT];  // CAP elements, one padding slot per thread
  for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint32_t bk = list ? list[q] : bucket0 + q;
    sort_one_bucket<W, CAP, NT>(bk, s, ik, ic, ok, oc, bstart, N, nb, bkb, big, lists, nlist, NBK, err);
    __syncthreads();  // s is reused by the next bucket
  }
}

// Oversized buckets to a contiguous array (dir 0: their records from the
// bucket sort's input) and back (dir 1: keys and codes to its output).
template <class SK>
__global__ void k_big_move(const uint64_t* __restrict__ bstart, const uint32_t* __restrict__ big,
                           const uint64_t* __restrict__ boff, const SK* __restrict__ ik, const uint8_t* __restrict__ ic,
                           SK* __restrict__ ok, uint8_t* __restrict__ oc, SK* __restrict__ tk,
                           uint8_t* __restrict__ tc, int dir, uint32_t bucket0, uint32_t bucket1) {
  for (uint32_t bk = bucket0 + blockIdx.x; bk < bucket1; bk += gridDim.x) {
    if (big[bk] != 1) continue;
    const uint64_t a = bstart[bk], n = bstart[bk + 1] - a, o = boff[bk];
    for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) {
      if (dir == 0) {
        tk[o + i] = ik[a + i];
        tc[o + i] = ic[a + i];
      } else {
        ok[a + i] = tk[o + i];
        oc[a + i] = tc[o + i];
      }
    }
  }
}

__global__ void k_big_sizes(uint32_t NB, const uint64_t* __restrict__ bstart, const uint32_t* __restrict__ big,
                            uint64_t* __restrict__ sz) {
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < NB; b += gridDim.x * blockDim.x)
    sz[b] = (big[b] == 1 && bstart[b + 1] >= bstart[b]) ? bstart[b + 1] - bstart[b] : 0;
}

// Nodes are the runs of equal keys in the sorted records.  Run starts per
// chunk of kEmitChunk records (a record starts a run when its key differs
// from the record before); their exclusive scan gives each chunk's first
// node, and k_node_emit numbers the chunk's records from it.  A block takes
// a chunk in kEmitRounds rounds of kBlock consecutive records, with every
// round's loads issued before the first barrier.
constexpr int kEmitRounds = 4;
constexpr int kEmitChunk = kBlock * kEmitRounds;

template <class SK>
__global__ void __launch_bounds__(kBlock) k_run_count(uint64_t N, const SK* __restrict__ keys,
                                                      uint32_t* __restrict__ cnt) {
  __shared__ uint32_t ws[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (uint64_t c = blockIdx.x; c * kEmitChunk < N; c += gridDim.x) {
    uint32_t t = 0;
#pragma unroll
    for (int r = 0; r < kEmitRounds; r++) {
      const uint64_t i = c * kEmitChunk + (uint64_t)r * kBlock + threadIdx.x;
      const bool f = i < N && (i == 0 || !sk_eq(keys[i], keys[i - 1]));
      t += (uint32_t)__popcll(__ballot(f));
    }
    if (lane == 0) ws[wid] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t a = 0;
      for (int w = 0; w < kBlock / 64; w++) a += ws[w];
      cnt[c] = a;
    }
    __syncthreads();
  }
}

// Run starts -> nodes (first record, key), and each node's neighbour mask:
// the OR of its records' code bits, gathered per block in LDS (a chunk's
// kEmitChunk records touch at most kEmitChunk consecutive nodes); the
// first and last node of a chunk may continue in a neighbouring chunk and
// are merged with a global atomicOr, the others stored directly.  Node of
// record i = the chunk's first node (coff, the scan of k_run_count) + the
// run starts in the chunk up to i - 1.  msk must start at zero.
template <int W>
__global__ void __launch_bounds__(kBlock) k_node_emit(uint64_t N, const typename SortKeyOf<W>::T* __restrict__ keys,
                                                      const uint8_t* __restrict__ rcode,
                                                      const uint64_t* __restrict__ coff, int nb,
                                                      uint64_t* __restrict__ klo, uint64_t* __restrict__ khi,
                                                      uint64_t* __restrict__ nstart, uint32_t* __restrict__ msk) {
  using SK = typename SortKeyOf<W>::T;
  __shared__ uint32_t lm[kEmitChunk];
  __shared__ uint32_t ws[kEmitRounds][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const uint64_t le = lane == 63 ? ~0ull : ((2ull << lane) - 1);  // lanes up to and including this one
  for (uint64_t c = blockIdx.x; c * kEmitChunk < N; c += gridDim.x) {
    const uint64_t base = c * kEmitChunk;
    SK key[kEmitRounds];
    uint32_t code[kEmitRounds];
    uint64_t bal[kEmitRounds];
#pragma unroll
    for (int r = 0; r < kEmitRounds; r++) {
      const uint64_t i = base + (uint64_t)r * kBlock + threadIdx.x;
      bool f = false;
      code[r] = 0;
      if (i < N) {
        key[r] = keys[i];
        code[r] = rcode[i];
        f = i == 0 || !sk_eq(key[r], keys[i - 1]);
      }
      bal[r] = __ballot(f);
      if (lane == 0) ws[r][wid] = (uint32_t)__popcll(bal[r]);
    }
    for (int q = threadIdx.x; q < kEmitChunk; q += kBlock) lm[q] = 0;
    const uint64_t c0 = coff[c], c1 = coff[c + 1];
    // the chunk's first node: its first record's run, which starts here or earlier
    const uint64_t nfirst = (base == 0 || !sk_eq(keys[base], keys[base - 1])) ? c0 : c0 - 1;
    __syncthreads();
    uint32_t before = 0;  // run starts of the earlier rounds
#pragma unroll
    for (int r = 0; r < kEmitRounds; r++) {
      const uint64_t i = base + (uint64_t)r * kBlock + threadIdx.x;
      uint32_t incl = before + (uint32_t)__popcll(bal[r] & le);
      for (int w = 0; w < kBlock / 64; w++) {
        if (w < wid) incl += ws[r][w];
        before += ws[r][w];
      }
      if (i < N) {
        const uint64_t n = c0 + incl - 1;
        if ((bal[r] >> lane) & 1) {
          nstart[n] = i;
          const KV v = key_of_sort(key[r], nb);
          klo[n] = v.lo;
          if (W == 2) khi[n] = v.hi;
        }
        atomicOr(&lm[n - nfirst], code_bits(code[r]));
      }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kEmitChunk; q += kBlock) {
      const uint64_t n = nfirst + q;
      if (n < c1) {
        const uint32_t bits = lm[q];
        if (n == nfirst || n == c1 - 1) atomicOr(&msk[n], bits);
        else msk[n] = bits;
      }
    }
    __syncthreads();
  }
}

__global__ void k_narrow_codes(uint64_t n, const uint32_t* __restrict__ in, uint8_t* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = (uint8_t)in[i];
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

__global__ void k_node_cnt(uint64_t nn, const uint64_t* __restrict__ nstart, uint32_t* __restrict__ cnt) {
  for (uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += (uint64_t)gridDim.x * blockDim.x)
    cnt[n] = (uint32_t)(nstart[n + 1] - nstart[n]);
}

// Key -> node lookup table (the unitig pass and the read heads find nodes
// by key): distinct keys, so a failed claim always means another key.
template <int W>
__global__ void k_node_table(uint64_t nn, const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi,
                             void* table, uint64_t cap, unsigned long long* err) {
  for (uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += (uint64_t)gridDim.x * blockDim.x) {
    KV key = node_key<W>(klo, khi, n);
    uint64_t s = slot_home(khash<W>(key), cap);
    bool done = false;
    for (uint64_t probe = 0; probe < cap && !done; probe++) {
      if (W == 1) {
        Slot1* T = (Slot1*)table;
        if (atomicCAS((unsigned long long*)&T[s].key, (unsigned long long)kEmpty, (unsigned long long)key.lo) ==
            kEmpty) {
          T[s].count = (uint32_t)n;
          done = true;
        }
      } else {
        Slot2* T = (Slot2*)table;
        if (atomicCAS((unsigned long long*)&T[s].hi, (unsigned long long)kEmpty, (unsigned long long)key.hi) ==
            kEmpty) {
          T[s].lo = key.lo;
          T[s].count = (uint32_t)n;
          done = true;
        }
      }
      s = slot_next(s, cap);
    }
    if (!done) report(err, ~0ull >> 8, ERR_TABLE);
  }
}

// Sorted node keys must be strictly increasing: one key in two table
// slots would break the claim protocol's guarantee; report the first.
template <int W>
__global__ void k_dup_check(uint64_t nn, const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi,
                            unsigned long long* err) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; i < nn; i += (uint64_t)gridDim.x * blockDim.x)
    if (klo[i] == klo[i - 1] && (W == 1 || khi[i] == khi[i - 1])) atomicMin(err, (unsigned long long)i);
}

template <int W>
__global__ void k_gather_nodes(uint64_t nn, const uint32_t* __restrict__ perm, const uint32_t* __restrict__ claimed,
                               void* table, uint64_t* __restrict__ klo, uint64_t* __restrict__ khi,
                               uint32_t* __restrict__ cnt, uint32_t* __restrict__ msk,
                               uint32_t* __restrict__ nslot) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t s = claimed[perm[i]];
    KV v = slot_key<W>(table, s);
    klo[i] = v.lo;
    if (W == 2) khi[i] = v.hi;
    uint32_t* c = W == 1 ? &((Slot1*)table)[s].count : &((Slot2*)table)[s].count;
    uint32_t* m = W == 1 ? &((Slot1*)table)[s].mask : &((Slot2*)table)[s].mask;
    cnt[i] = *c;
    msk[i] = *m;
    *c = (uint32_t)i;
    nslot[i] = s;
  }
}


__device__ inline void type_sizes(uint32_t mask, int n[4]) {
  for (int t = 0; t < 4; t++) n[t] = __popc((mask >> (4 * t)) & 0xF);
}

// Nodes whose lists need the replay: some list holds >= 2 k-mers.  Marks
// the slot (kNeedsBit), counts the records to replay and emits the replay
// list key (count, clamped) for length-balanced scheduling.
template <int W>
__global__ void k_node_info(uint64_t nn, const uint32_t* __restrict__ msk, const uint32_t* __restrict__ cnt,
                            const uint32_t* __restrict__ nslot, void* table, int jdk7,
                            uint32_t* __restrict__ needs, uint64_t* __restrict__ selcnt,
                            unsigned long long* maxcnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    if (cnt[i] > 255) atomicMax(maxcnt, (unsigned long long)cnt[i]);
    int n[4];
    type_sizes(msk[i], n);
    bool need = jdk7 && (n[0] > 1 || n[1] > 1 || n[2] > 1 || n[3] > 1);
    needs[i] = need ? 1 : 0;
    selcnt[i] = need ? cnt[i] : 0;
    if (need && table) {
      uint32_t s = nslot[i];
      uint32_t* m = W == 1 ? &((Slot1*)table)[s].mask : &((Slot2*)table)[s].mask;
      *m |= kNeedsBit;
    }
  }
}

// Stored neighbour k-mer of list t, letter x.
template <int W>
__device__ inline KV neighbour(KV v, int k, uint32_t t, uint32_t x) {
  switch (t) {
    case ET_FF: return succ_of<W>(v, k, x);
    case ET_FR: return revcomp<W>(succ_of<W>(v, k, x), k);
    case ET_RF: return revcomp<W>(pred_of<W>(v, k, x), k);
    default: return pred_of<W>(v, k, x);
  }
}

// ---------------------------------------------------------------------------
// 6. JDK-7 edge order.  Every record of a node passes through
//    VKmerList.unionUpdate (VKmerList.java:117-133) per edge type: a new
//    java.util.HashSet(|L| + |rec|) receives L then the record's k-mer and
//    the list becomes its iteration order; the cross aggregator repeats it
//    once more with an empty record (JobGenBuildBrujinGraph.java:119-143,
//    P = 1).  Records are visited in emission order, which the stable sort of
//    the reference keeps (FrameSorterMergeSort.java:153-185).
// ---------------------------------------------------------------------------

template <int W>
__global__ void k_sel_flags(uint64_t nocc, const uint32_t* __restrict__ occ_slot, const void* table,
                            uint32_t* __restrict__ flag) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nocc; j += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t s = occ_slot[j];
    uint32_t m = W == 1 ? ((const Slot1*)table)[s].mask : ((const Slot2*)table)[s].mask;
    flag[j] = m >> 31;
  }
}

template <int W>
__global__ void k_sel_write(uint64_t nocc, const uint32_t* __restrict__ occ_slot, const uint8_t* __restrict__ occ_code,
                            const void* table, const uint32_t* __restrict__ flag, const uint64_t* __restrict__ pos,
                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nocc; j += (uint64_t)gridDim.x * blockDim.x) {
    if (flag[j]) {
      uint64_t p = pos[j];
      uint32_t s = occ_slot[j];
      keys[p] = W == 1 ? ((const Slot1*)table)[s].count : ((const Slot2*)table)[s].count;
      vals[p] = occ_code[j];
    }
  }
}

// Replay list: nodes needing the replay, keyed by descending record count
// so that one wave gets nodes of similar length.
__global__ void k_replay_list(uint64_t nn, const uint32_t* __restrict__ needs, const uint64_t* __restrict__ rpos,
                              const uint32_t* __restrict__ cnt, uint32_t* __restrict__ rkey,
                              uint32_t* __restrict__ rval) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    if (needs[i]) {
      uint64_t p = rpos[i];
      uint32_t c = cnt[i] > 0xffffu ? 0xffffu : cnt[i];
      rkey[p] = 0xffffu - c;
      rval[p] = (uint32_t)i;
    }
  }
}

// Small ordered list of 2-bit letters (iteration order of the HashSet).
struct LList {
  uint32_t bits;  // letter j at bits 2j
  int n;
};

__device__ inline uint32_t ll_get(const LList& l, int j) { return (l.bits >> (2 * j)) & 3; }

// One VKmerList.unionUpdate with at most one new element.  The lists hold
// at most 4 k-mers, so the HashSet capacity never exceeds 8 (capacity
// nextPow2(<= 5); a resize needs size >= 0.75 * capacity with a collision,
// which with <= 4 elements only happens at capacity 4) and only the low 3
// bits of each element's JDK-7-spread hash matter: hb packs them, letter x
// at bits 3x..3x+2.
__device__ inline uint32_t bucket_of(uint32_t hb, uint32_t x, int cap) { return (hb >> (3 * x)) & (uint32_t)(cap - 1); }

__device__ inline LList j7_union(LList L, int sym, uint32_t hb) {
  int n = L.n + (sym >= 0 ? 1 : 0);
  int cap = n <= 1 ? 1 : (n <= 2 ? 2 : (n <= 4 ? 4 : 8));
  int thr = (cap * 3) >> 2;  // (int)(cap * 0.75f)
  LList o{0, 0};
  for (int idx = 0; idx < n; idx++) {
    uint32_t x = idx < L.n ? ll_get(L, idx) : (uint32_t)sym;
    bool present = false;
    uint32_t b = bucket_of(hb, x, cap);
    bool nonempty = false;
    for (int j = 0; j < o.n; j++) {
      uint32_t y = ll_get(o, j);
      present |= y == x;
      nonempty |= bucket_of(hb, y, cap) == b;
    }
    if (present) continue;
    if (o.n >= thr && nonempty) {
      // HashMap.resize/transfer: new order = by new bucket, ties in reverse
      // of the old iteration order.
      int ncap = cap * 2;
      LList r{0, 0};
      for (int bk = 0; bk < ncap; bk++)
        for (int j = o.n - 1; j >= 0; j--) {
          uint32_t y = ll_get(o, j);
          if (bucket_of(hb, y, ncap) == (uint32_t)bk) { r.bits |= y << (2 * r.n); r.n++; }
        }
      o = r;
      cap = ncap;
      thr = (cap * 3) >> 2;
      b = bucket_of(hb, x, cap);
    }
    // head of bucket b: after every element of a lower bucket
    int pos = 0;
    for (int j = 0; j < o.n; j++) pos += bucket_of(hb, ll_get(o, j), cap) < b ? 1 : 0;
    uint32_t low = o.bits & ((1u << (2 * pos)) - 1);
    uint32_t high = o.bits >> (2 * pos);
    o.bits = low | (x << (2 * pos)) | (high << (2 * pos + 2));
    o.n++;
  }
  return o;
}

// Long runs (a repeat k-mer carries thousands of records) are replayed by
// four lanes per node, one per edge type: the four lists are independent
// state machines over the same records, so the serial chain per lane is a
// quarter of the work.  Lanes 4q..4q+3 hold node q; the four 16-bit results
// are merged with lane shuffles.  Each lane caches its last eight
// transitions (n, order, symbol class) -> (n', order'): a list of fixed
// letters moves between a handful of orders under the two reorderings (a
// union with no letter, with a letter present), so after a few records every
// lane hits its cache and the wave no longer runs the full union whenever
// one of its 64 lanes misses.
constexpr uint32_t kLongReplay = 64;

__device__ inline int code_sym(uint32_t code, int t) {
  uint32_t sn = code & 0xF, pn = (code >> 4) & 0xF;
  if (t == ET_FF || t == ET_FR) return ((sn & 8) && (((sn & 4) != 0) == (t == ET_FR))) ? (int)(sn & 3) : -1;
  return ((pn & 8) && (((pn & 4) != 0) == (t == ET_RF))) ? (int)(pn & 3) : -1;
}

template <int W>
__global__ void k_replay_long(uint64_t nlong, const uint32_t* __restrict__ rlist, const uint64_t* __restrict__ seg,
                              const uint8_t* __restrict__ codes, const uint64_t* __restrict__ klo,
                              const uint64_t* __restrict__ khi, const uint32_t* __restrict__ msk, int k, int nb,
                              uint64_t* __restrict__ lists) {
  for (uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; gid < 4 * nlong;
       gid += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t q = gid >> 2;
    const int t = (int)(gid & 3);
    const uint32_t i = rlist[q];
    const uint32_t mask = msk[i];
    const bool active = __popc((mask >> (4 * t)) & 0xF) > 1;
    uint64_t part = 0;
    if (active) {
      KV v = node_key<W>(klo, khi, i);
      uint32_t hb = 0;
      for (uint32_t x = 0; x < 4; x++)
        hb |= ((uint32_t)jdk7_spread(java_hash_vkmer<W>(neighbour<W>(v, k, (uint32_t)t, x), k, nb)) & 7u) << (3 * x);
      LList l{0, 0};
      uint32_t memo[8];  // key << 11 | (n' << 8 | order'); all ones = empty (keys have 14 bits)
#pragma unroll
      for (int m = 0; m < 8; m++) memo[m] = 0xffffffffu;
      uint32_t rr = 0;  // slot the next miss fills
      const uint64_t a = seg[i], e = seg[i + 1];
      uint64_t g = a & ~15ull;
      uint4 cur = *reinterpret_cast<const uint4*>(codes + g);
      while (g < e) {
        const uint64_t gn = g + 16;
        const uint4 nxt = gn < e ? *reinterpret_cast<const uint4*>(codes + gn) : make_uint4(0, 0, 0, 0);
        const uint64_t lo64 = (uint64_t)cur.x | ((uint64_t)cur.y << 32), hi64 = (uint64_t)cur.z | ((uint64_t)cur.w << 32);
        const int j0 = g < a ? (int)(a - g) : 0, j1 = e - g < 16 ? (int)(e - g) : 16;
        for (int j = j0; j < j1; j++) {
          uint32_t code = (uint32_t)((j < 8 ? lo64 >> (8 * j) : hi64 >> (8 * (j - 8))) & 0xff);
          int sy = code_sym(code, t);
          if (l.n <= 1 && (sy < 0 || (l.n == 1 && (int)ll_get(l, 0) == sy))) continue;
          if (l.n == 0) { l.bits = (uint32_t)sy; l.n = 1; continue; }
          uint32_t cls = 5;
          if (sy >= 0) {
            cls = (uint32_t)sy;
            for (int jj = 0; jj < l.n; jj++)
              if ((int)ll_get(l, jj) == sy) cls = 4;
          }
          uint32_t key = ((uint32_t)l.n << 11) | (l.bits << 3) | cls;
          uint32_t hit = 0xffffffffu;
#pragma unroll
          for (int m = 0; m < 8; m++) hit = (memo[m] >> 11) == key ? memo[m] : hit;
          if (hit != 0xffffffffu) { l.n = (hit >> 8) & 7; l.bits = hit & 0xff; continue; }
          LList o = j7_union(l, sy, hb);
          const uint32_t nv = (key << 11) | ((uint32_t)o.n << 8) | o.bits;
#pragma unroll
          for (int m = 0; m < 8; m++) memo[m] = m == (int)rr ? nv : memo[m];
          rr = (rr + 1) & 7;
          l = o;
        }
        cur = nxt;
        g = gn;
      }
      LList f = j7_union(l, -1, hb);  // cross aggregator init
      part = ((uint64_t)f.n | ((uint64_t)f.bits << 3)) << (16 * t);
    }
    uint32_t plo = (uint32_t)part, phi = (uint32_t)(part >> 32);
    plo |= __shfl_xor(plo, 1);
    phi |= __shfl_xor(phi, 1);
    plo |= __shfl_xor(plo, 2);
    phi |= __shfl_xor(phi, 2);
    if (t == 0) lists[i] = ((uint64_t)phi << 32) | plo;
  }
}

__global__ void k_count_long(uint64_t n, const uint32_t* __restrict__ rkey, uint32_t lim,
                             unsigned long long* __restrict__ cnt) {
  uint64_t c = 0;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (uint64_t)gridDim.x * blockDim.x)
    c += rkey[q] <= lim;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(cnt, (unsigned long long)c);  // one add per wave
}

// One thread per node of the replay list.  Only lists that end with >= 2
// k-mers are simulated; a list with <= 1 k-mer has one order.  Fast paths:
// unions that cannot reorder (empty or one-element list, no new element)
// are skipped, and a union's result depends only on (list size, order,
// symbol class) -- a new letter, a letter already present, or none -- so
// the last two transitions of each list are cached.
template <int W>
__global__ void k_replay(uint64_t nrep, const uint32_t* __restrict__ rlist, const uint64_t* __restrict__ seg,
                         const uint8_t* __restrict__ codes, const uint64_t* __restrict__ klo,
                         const uint64_t* __restrict__ khi, const uint32_t* __restrict__ msk, int k, int nb,
                         uint64_t* __restrict__ lists) {
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nrep; q += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t i = rlist[q];
    KV v = node_key<W>(klo, khi, i);
    uint32_t mask = msk[i];
    uint32_t rel = 0;
    int nfin[4];
    for (int t = 0; t < 4; t++) {
      nfin[t] = __popc((mask >> (4 * t)) & 0xF);
      if (nfin[t] > 1) rel |= 1u << t;
    }
    // low 3 bits of the JDK-7-spread VKmer.hashCode of each stored neighbour
    uint32_t hb[4] = {0, 0, 0, 0};
    for (uint32_t t = 0; t < 4; t++)
      if ((rel >> t) & 1)
        for (uint32_t x = 0; x < 4; x++)
          hb[t] |= ((uint32_t)jdk7_spread(java_hash_vkmer<W>(neighbour<W>(v, k, t, x), k, nb)) & 7u) << (3 * x);
    LList L[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    // two-entry transition cache per list: (n, order, symbol class) ->
    // (n', order'); an all-ones word is empty (no valid key has n = 7)
    uint32_t cache[4][2];
    for (int t = 0; t < 4; t++) cache[t][0] = cache[t][1] = 0xffffffffu;
    uint64_t a = seg[i], e = seg[i + 1];
    // Codes are bytes, read 16 at a time with aligned 16-byte loads; the
    // next group is loaded before the current one is replayed, so a long
    // list (a repeat k-mer has thousands of records) waits on memory only
    // once per group.  The code array is padded to a multiple of 16 bytes.
    uint64_t g = a & ~15ull;
    uint4 cur = a < e ? *reinterpret_cast<const uint4*>(codes + g) : make_uint4(0, 0, 0, 0);
    while (g < e) {
      const uint64_t gn = g + 16;
      const uint4 nxt = gn < e ? *reinterpret_cast<const uint4*>(codes + gn) : make_uint4(0, 0, 0, 0);
      const uint64_t lo64 = (uint64_t)cur.x | ((uint64_t)cur.y << 32), hi64 = (uint64_t)cur.z | ((uint64_t)cur.w << 32);
      const int j0 = g < a ? (int)(a - g) : 0, j1 = e - g < 16 ? (int)(e - g) : 16;
      for (int j = j0; j < j1; j++) {
        uint32_t code = (uint32_t)((j < 8 ? lo64 >> (8 * j) : hi64 >> (8 * (j - 8))) & 0xff);
        int sym[4] = {-1, -1, -1, -1};
        uint32_t sn = code & 0xF, pn = (code >> 4) & 0xF;
        if (sn & 8) sym[(sn & 4) ? ET_FR : ET_FF] = (int)(sn & 3);
        if (pn & 8) sym[(pn & 4) ? ET_RF : ET_RR] = (int)(pn & 3);
        for (int t = 0; t < 4; t++) {
          if (!((rel >> t) & 1)) continue;
          int sy = sym[t];
          LList& l = L[t];
          if (l.n <= 1 && (sy < 0 || (l.n == 1 && (int)ll_get(l, 0) == sy))) continue;
          if (l.n == 0) { l.bits = (uint32_t)sy; l.n = 1; continue; }
          // symbol class: the new letter, 4 = a letter already present, 5 = none
          uint32_t cls = 5;
          if (sy >= 0) {
            cls = (uint32_t)sy;
            for (int j = 0; j < l.n; j++)
              if ((int)ll_get(l, j) == sy) cls = 4;
          }
          uint32_t key = ((uint32_t)l.n << 11) | (l.bits << 3) | cls;  // 14 bits
          uint32_t c0 = cache[t][0], c1 = cache[t][1];
          if ((c0 >> 11) == key) { l.n = (c0 >> 8) & 7; l.bits = c0 & 0xff; continue; }
          if ((c1 >> 11) == key) { l.n = (c1 >> 8) & 7; l.bits = c1 & 0xff; continue; }
          LList o = j7_union(l, sy, hb[t]);
          cache[t][1] = c0;
          cache[t][0] = (key << 11) | ((uint32_t)o.n << 8) | o.bits;
          l = o;
        }
      }
      cur = nxt;
      g = gn;
    }
    uint64_t out = 0;
    for (int t = 0; t < 4; t++) {
      if (!((rel >> t) & 1)) continue;
      LList f = j7_union(L[t], -1, hb[t]);  // cross aggregator init
      out |= ((uint64_t)f.n | ((uint64_t)f.bits << 3)) << (16 * t);
    }
    lists[i] = out;
  }
}

// ---------------------------------------------------------------------------
// 7. 5' read heads (ReadHeadInfo.java:100-136; ReadsKeyValueParserFactory.java
//    :129-135,142-144,165-170): one per valid mate on its first k-mer,
//    unflipped at offset 0 or flipped at offset k-1.  Sort key =
//    node<<31 | flipped<<30 | mate<<29 | readId: within a set the TreeSet
//    order (offset, lib = 0, mate, readId) of ReadHeadInfo.compareTo
//    (:247-264); equal keys keep the first in emission order (TreeSet.add).
// ---------------------------------------------------------------------------

template <int W>
__global__ void k_heads(uint64_t nmates, const uint8_t* __restrict__ valid, const uint64_t* __restrict__ nocc,
                        const uint64_t* __restrict__ hidx, NodeIndex ix,
                        const int64_t* __restrict__ rid, const uint64_t* __restrict__ wbase,
                        const uint64_t* __restrict__ words, int k, uint64_t* __restrict__ hkey,
                        uint32_t* __restrict__ hval, uint32_t tlo, uint32_t thi) {
  using SK = typename SortKeyOf<W>::T;
  for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < nmates; m += (uint64_t)gridDim.x * blockDim.x) {
    if (!valid[m] || nocc[m] == 0) continue;
    bool fwd;
    KV kk = canon<W>(load_kmer<W>(words + wbase[m], 0, k), k, &fwd);
    if (tlo != 0 || thi != 256) {
      SK sk;
      make_sort_key(kk, ix.nb, sk);
      const uint32_t t = key_byte(sk, ix.nb - 1);
      if (t < tlo || t >= thi) {  // another key range holds this head's node
        hkey[hidx[m]] = ~0ull;
        hval[hidx[m]] = 0;
        continue;
      }
    }
    uint64_t node = node_lookup<W>(ix, kk);
    uint64_t flipped = fwd ? 0 : 1;
    uint64_t key = (node << 31) | (flipped << 30) | ((m & 1) << 29) | (uint64_t)rid[m >> 1];
    uint64_t j = hidx[m];
    hkey[j] = key;
    hval[j] = (uint32_t)m;
  }
}

// Heads of the current key range (the others carry key ~0), kept in mate
// order by a scan of these flags.
__global__ void k_head_in_range(uint64_t nh, const uint64_t* __restrict__ hkey, uint32_t* __restrict__ flag) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += (uint64_t)gridDim.x * blockDim.x)
    flag[j] = hkey[j] != ~0ull ? 1u : 0u;
}

__global__ void k_head_keep(uint64_t nh, const uint64_t* __restrict__ hkey, const uint32_t* __restrict__ hval,
                            const uint32_t* __restrict__ flag, const uint64_t* __restrict__ pos,
                            uint64_t* __restrict__ okey, uint32_t* __restrict__ oval) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += (uint64_t)gridDim.x * blockDim.x)
    if (flag[j]) {
      okey[pos[j]] = hkey[j];
      oval[pos[j]] = hval[j];
    }
}

// Unique heads (TreeSet keeps the first of equal keys) and per-node counts
// of unflipped / flipped heads.
__global__ void k_head_unique(uint64_t nh, const uint64_t* __restrict__ hkey, uint32_t* __restrict__ uflag) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += (uint64_t)gridDim.x * blockDim.x)
    uflag[j] = hkey[j] != ~0ull && (j == 0 || hkey[j] != hkey[j - 1]) ? 1 : 0;  // ~0: head of another range
}

__global__ void k_head_compact(uint64_t nh, const uint64_t* __restrict__ hkey, const uint32_t* __restrict__ hval,
                               const uint32_t* __restrict__ uflag, const uint64_t* __restrict__ upos,
                               uint32_t* __restrict__ uval, uint32_t* __restrict__ hcu, uint32_t* __restrict__ hcf) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += (uint64_t)gridDim.x * blockDim.x) {
    if (!uflag[j]) continue;
    uint64_t key = hkey[j];
    uval[upos[j]] = hval[j] | ((uint32_t)((key >> 30) & 1) << 31);
    uint64_t node = key >> 31;
    atomicAdd(((key >> 30) & 1) ? &hcf[node] : &hcu[node], 1u);
  }
}

__global__ void k_head_total(uint64_t nn, const uint32_t* __restrict__ hcu, const uint32_t* __restrict__ hcf,
                             uint64_t* __restrict__ tot) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x)
    tot[i] = (uint64_t)hcu[i] + hcf[i];
}

// ---------------------------------------------------------------------------
// 8. Node serialisation (Node.write, Node.java:408-487; VKmerList.write
//    :314-316; ExternalableTreeSet.write with writeEntireBody :236-267;
//    ReadHeadInfo.write :205-212; VKmer.write :389-391) inside a SequenceFile
//    record (i32 recLen, i32 keyLen, VKmer key, Node).
// ---------------------------------------------------------------------------

struct SerCtx {
  int k, nb, jdk7;
  uint64_t nn;
  const uint64_t *klo, *khi;
  const uint32_t *cnt, *msk;
  const uint32_t* needs;
  const uint64_t* lists;
  const uint64_t* hstart;  // first unique head of each node
  const uint32_t *hcu, *hcf;
  const uint32_t* uval;     // unique heads: mate index | flipped << 31
  const uint32_t* mlen;
  const uint64_t* wbase;
  const uint64_t* words;
  const int64_t* rid;
  unsigned long long* bad;  // internal consistency flag (first bad node)
  const uint32_t* perm;     // output position -> node (partitioned output), or null
};

__device__ inline uint64_t ser_node(const SerCtx& c, uint64_t j) { return c.perm ? c.perm[j] : j; }

__device__ inline uint32_t nbytes_of(uint32_t L) { return (L + 3) / 4; }

__device__ inline uint64_t head_bytes(const SerCtx& c, uint32_t m) {
  uint32_t L = c.mlen[m], Lm = c.mlen[m ^ 1];
  return 1 + 8 + 4 + nbytes_of(L) + (Lm > 0 ? 4 + nbytes_of(Lm) : 0);
}

// Letters of edge type t in output order, from the node's mask, replay
// flag and replayed lists (read once per node by the caller).
template <int W>
__device__ inline LList edge_list(const SerCtx& c, uint32_t mask, uint32_t needs, uint64_t lists, int t, KV v) {
  uint32_t sel = (mask >> (4 * t)) & 0xF;
  LList l{0, 0};
  if (c.jdk7 && needs && __popc(sel) > 1) {
    uint64_t w = lists >> (16 * t);
    l.n = (int)(w & 7);
    l.bits = (uint32_t)(w >> 3) & 0xFF;
    return l;
  }
  // ascending neighbour k-mer (numeric V = VKmer byte order): the rank of
  // each letter is the number of smaller neighbours (the neighbours of one
  // type are distinct); fully unrolled so everything stays in registers
  KV nk[4];
#pragma unroll
  for (int x = 0; x < 4; x++) nk[x] = ((sel >> x) & 1) ? neighbour<W>(v, c.k, (uint32_t)t, (uint32_t)x) : KV{0, 0};
#pragma unroll
  for (int x = 0; x < 4; x++) {
    if (!((sel >> x) & 1)) continue;
    int r = 0;
#pragma unroll
    for (int y = 0; y < 4; y++)
      if (y != x && ((sel >> y) & 1) && kle<W>(nk[y], nk[x])) r++;
    l.bits |= (uint32_t)x << (2 * r);
  }
  l.n = __popc(sel);
  return l;
}

__device__ inline uint64_t record_size(const SerCtx& c, uint64_t i) {
  int n[4];
  type_sizes(c.msk[i], n);
  uint64_t val = 1 + 4;
  for (int t = 0; t < 4; t++)
    if (n[t]) val += 4 + (uint64_t)n[t] * (4 + c.nb);
  uint64_t h0 = c.hstart[i];
  uint32_t nu = c.hcu[i], nf = c.hcf[i];
  if (nu) {
    val += 1 + 4;
    for (uint32_t j = 0; j < nu; j++) val += head_bytes(c, c.uval[h0 + j] & 0x7fffffffu);
  }
  if (nf) {
    val += 1 + 4;
    for (uint32_t j = 0; j < nf; j++) val += head_bytes(c, c.uval[h0 + nu + j] & 0x7fffffffu);
  }
  return 8 + 4 + c.nb + val;
}

__global__ void k_ser_size(SerCtx c, uint64_t* __restrict__ rsize) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.nn; i += (uint64_t)gridDim.x * blockDim.x)
    rsize[i] = record_size(c, ser_node(c, i));
}

// Record bytes are written with 2-, 4- and 8-byte stores at any alignment
// (gfx950 runs in unaligned access mode; the compiler merges these copies
// into single stores), never past the bytes asked for.

// The n (<= 8) low bytes of x, least significant first.
__device__ inline void put_le(uint8_t*& p, uint64_t x, int n) {
  if (n == 8) {
    memcpy(p, &x, 8);
    p += 8;
    return;
  }
  if (n & 4) {
    uint32_t y = (uint32_t)x;
    memcpy(p, &y, 4);
    p += 4;
    x >>= 32;
  }
  if (n & 2) {
    uint16_t y = (uint16_t)x;
    memcpy(p, &y, 2);
    p += 2;
    x >>= 16;
  }
  if (n & 1) *p++ = (uint8_t)x;
}

__device__ inline void put32(uint8_t*& p, uint32_t v) {  // big endian
  uint32_t b = __builtin_bswap32(v);
  memcpy(p, &b, 4);
  p += 4;
}

__device__ inline void put64(uint8_t*& p, uint64_t v) {  // big endian
  uint64_t b = __builtin_bswap64(v);
  memcpy(p, &b, 8);
  p += 8;
}

// VKmer bytes: BE32(k), then the nb-byte big-endian form of V (whose
// little-endian image is the sort key).
template <int W>
__device__ inline void put_key(uint8_t*& p, KV v, int k, int nb) {
  put32(p, (uint32_t)k);
  if constexpr (W == 1) {
    put_le(p, sortkey64(v.lo, nb), nb);
  } else {
    Key128 s = sortkey128(v, nb);
    put_le(p, s.lo, nb < 8 ? nb : 8);
    if (nb > 8) put_le(p, s.hi, nb - 8);
  }
}

// A read as VKmer bytes: BE32(L) then the packed stream's bytes reversed,
// eight at a time as byte-swapped 64-bit windows of the stream.
__device__ inline void put_read(uint8_t*& p, const uint64_t* w, uint32_t L) {
  put32(p, L);
  const int64_t nbl = nbytes_of(L);
  int64_t q = nbl - 8;  // lowest stream byte of the next window
  for (; q >= 0; q -= 8) {
    const uint64_t a = (uint64_t)q >> 3, sh = ((uint64_t)q & 7) * 8;
    uint64_t x = sh ? (w[a] >> sh) | (w[a + 1] << (64 - sh)) : w[a];
    put64(p, x);
  }
  for (int64_t r = q + 7; r >= 0; r--) *p++ = (uint8_t)(w[r >> 3] >> (8 * (r & 7)));
}

template <int W>
__device__ inline void write_record(const SerCtx& c, uint64_t i, uint8_t* p, uint64_t rec) {
  // every per-node field is loaded here, before the first store: the
  // stores through p could alias them, which would order each later load
  // after them
  const KV v = node_key<W>(c.klo, c.khi, i);
  const uint32_t mask = c.msk[i], nu = c.hcu[i], nf = c.hcf[i], cov = c.cnt[i];
  const uint32_t needs = c.jdk7 ? c.needs[i] : 0u;
  const uint64_t lists = needs ? c.lists[i] : 0ull;
  const uint64_t h0 = (nu | nf) ? c.hstart[i] : 0ull;
  put32(p, (uint32_t)(rec - 8));
  put32(p, 4 + c.nb);
  put_key<W>(p, v, c.k, c.nb);
  int n[4];
  type_sizes(mask, n);
  uint8_t flags = 0x80;
  for (int t = 0; t < 4; t++)
    if (n[t]) flags |= 1 << t;
  if (nu) flags |= 1 << 4;
  if (nf) flags |= 1 << 5;
  *p++ = flags;
  for (int t = 0; t < 4; t++) {
    if (!n[t]) continue;
    LList l = edge_list<W>(c, mask, needs, lists, t, v);
    if (l.n != n[t]) {  // cannot happen: the replayed set holds every neighbour once
      if (c.bad) atomicMin(c.bad, (unsigned long long)i);
      l.n = n[t];
    }
    put32(p, (uint32_t)n[t]);
    for (int a = 0; a < n[t]; a++) put_key<W>(p, neighbour<W>(v, c.k, (uint32_t)t, ll_get(l, a)), c.k, c.nb);
  }
  for (int f = 0; f < 2; f++) {
    uint32_t cntf = f ? nf : nu;
    if (!cntf) continue;
    uint64_t a = h0 + (f ? nu : 0);
    *p++ = 1;  // whole body in stream
    put32(p, cntf);
    for (uint64_t j = a; j < a + cntf; j++) {
      uint32_t m = c.uval[j] & 0x7fffffffu;
      uint32_t L = c.mlen[m], Lm = c.mlen[m ^ 1];
      *p++ = Lm > 0 ? 1 : 0;
      uint64_t off = f ? (uint64_t)(c.k - 1) : 0;
      uint64_t uuid = (off << 40) | ((uint64_t)(m & 1) << 35) | (uint64_t)c.rid[m >> 1];
      put64(p, uuid);
      put_read(p, c.words + c.wbase[m], L);
      if (Lm > 0) put_read(p, c.words + c.wbase[m ^ 1], Lm);
    }
  }
  // coverage: Java float accumulation of 1.0f saturates at 2^24
  uint32_t cv = cov > (1u << 24) ? (1u << 24) : cov;
  float fc = (float)cv;
  uint32_t u;
  memcpy(&u, &fc, 4);
  put32(p, u);
}

// Records of 256 consecutive nodes are contiguous in the output: each
// thread writes its record into LDS, then the block stores the whole span
// with coalesced 4-byte writes.  Spans larger than the staging buffer are
// written directly.
constexpr int kSerStage = 16384;  // 256 node records average ~60 bytes; a smaller stage keeps 8 waves per SIMD

template <int W>
__global__ void __launch_bounds__(kBlock) k_ser_write(SerCtx c, const uint64_t* __restrict__ roff,
                                                      uint8_t* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[kSerStage + 8];
  uint64_t first = (uint64_t)blockIdx.x * kBlock;
  if (first >= c.nn) return;
  uint64_t last = first + kBlock < c.nn ? first + kBlock : c.nn;
  uint64_t s0 = roff[first], s1 = roff[last];
  uint64_t i = first + threadIdx.x;
  // the stage starts at the 4-byte boundary at or below s0
  uint64_t base = s0 & ~3ull;
  bool staged = (s1 - base) <= kSerStage;
  if (i < last) {
    uint64_t rec = roff[i + 1] - roff[i];
    write_record<W>(c, ser_node(c, i), staged ? stage + (roff[i] - base) : out + roff[i], rec);
  }
  if (!staged) return;
  __syncthreads();
  // head and tail bytes that are not 4-byte aligned
  uint64_t a4 = (s0 + 3) & ~3ull, e4 = s1 & ~3ull;
  if (a4 > e4) { a4 = s1; e4 = s1; }
  for (uint64_t x = s0 + threadIdx.x; x < a4; x += kBlock) out[x] = stage[x - base];
  for (uint64_t x = e4 + threadIdx.x; x < s1; x += kBlock) out[x] = stage[x - base];
  const uint32_t* src = reinterpret_cast<const uint32_t*>(stage + (a4 - base));
  uint32_t* dst = reinterpret_cast<uint32_t*>(out + a4);
  for (uint64_t x = threadIdx.x; x < (e4 - a4) / 4; x += kBlock)
    __builtin_nontemporal_store(src[x], dst + x);  // written once, read by the caller later
}

// Output partition of each node: nonneg(hashBytes(key)) % P, the reference's
// KmerPartitionComputerFactory (KmerPartitionComputerFactory.java:28-53),
// where nonneg(h) = h < 0 ? -(h + 1) : h.
template <int W>
__global__ void k_part_ids(uint64_t nn, const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi, int nb,
                           uint32_t P, uint32_t* __restrict__ pid, uint32_t* __restrict__ idx) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    int32_t h = java_hash_key<W>(node_key<W>(klo, khi, i), nb);
    if (h < 0) h = -(h + 1);
    pid[i] = (uint32_t)h % P;
    idx[i] = (uint32_t)i;
  }
}

__global__ void k_part_bytes(uint32_t P, const uint64_t* __restrict__ start, const uint64_t* __restrict__ roff,
                             uint64_t* __restrict__ boff) {
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p <= P; p += gridDim.x * blockDim.x) boff[p] = roff[start[p]];
}

// First output position of every partition, from the partition ids sorted
// stably (start[P] = nn; empty partitions start where the next one does).
__global__ void k_part_starts(uint64_t nn, const uint32_t* __restrict__ spid, uint32_t P,
                              uint64_t* __restrict__ start) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= nn; j += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = j == 0 ? 0 : spid[j - 1] + 1;
    uint32_t hi = j == nn ? P : spid[j];
    for (uint32_t p = lo; p <= hi && p <= P; p++) start[p] = j;
  }
}

// words per mate for the packed store
__global__ void k_word_counts(uint64_t nmates, const uint32_t* __restrict__ mlen, uint64_t* __restrict__ wc) {
  for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < nmates; m += (uint64_t)gridDim.x * blockDim.x)
    wc[m] = (mlen[m] + 31) / 32 + 1;  // +1: k-mer loads may touch one word past the end
}

__global__ void k_valid_u64(uint64_t nmates, const uint8_t* __restrict__ valid, const uint64_t* __restrict__ nocc,
                            uint64_t* __restrict__ out) {
  for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < nmates; m += (uint64_t)gridDim.x * blockDim.x)
    out[m] = (valid[m] && nocc[m]) ? 1 : 0;
}

// ===========================================================================
// Unitig compression (MERGE / P4 path merge) on the built graph.
//
// The fixed point of P4ForPathMergeVertex (P4ForPathMergeVertex.java:21-266;
// BasicPathMergeVertex.java:33-219) merges exactly the links X -(t)- Y where
// X's side of t and Y's side of mirror(t) both have degree 1, X != Y and
// neither node lists itself (a tandem repeat restricts both directions,
// DeBruijnGraphCleanVertex.java:234-245).  Here the merged paths are found by
// pointer-jumping list ranking over node ends: element 2i+d traverses node i
// leaving through side d (0 = F, 1 = R); its successor is the element of
// the merge partner that leaves through the partner's other side.  The
// surviving k-mer (random in P4) is the path's first k-mer in KmerPointable
// order, the path is oriented so that it reads forward, and a cycle is cut
// at the link entering that k-mer.
// ===========================================================================

constexpr int kUEdges = 16;
// A unitig edge entry is (type << kRepBits) | the neighbour unitig's
// representative node, so the unitig pass, the cleaning steps and BUBBLE
// accept graphs of fewer than kMaxUnitigNodes nodes.
constexpr int kRepBits = 30;
constexpr uint32_t kRepMask = (1u << kRepBits) - 1;
constexpr uint64_t kMaxUnitigNodes = 1ull << kRepBits;
// A read head's letter offset in a unitig has ReadHeadInfo's 23 bits
// (ReadHeadInfo.java:20,34; makeUUID throws beyond them, :115-119).
constexpr uint64_t kMaxHeadOffset = (1ull << 23) - 1;
// The largest representative and every type survive the packing; the sort
// keys of unitig read heads (flipped << 53 | offset << 30 | mate << 29 |
// readId) keep the offset field clear of the flag.
static_assert((((3u << kRepBits) | (uint32_t)(kMaxUnitigNodes - 1)) >> kRepBits) == 3u, "edge type field");
static_assert((((3u << kRepBits) | (uint32_t)(kMaxUnitigNodes - 1)) & kRepMask) == kMaxUnitigNodes - 1,
              "edge representative field");
static_assert(((kMaxHeadOffset << 30) >> 53) == 0, "head offset below the flipped bit");

__device__ inline int side_deg(uint32_t mask, int d) {
  return d == 0 ? __popc(mask & 0xFFu) : __popc((mask >> 8) & 0xFFu);
}
// EDGETYPE.mirror / dir (EDGETYPE.java:44-60,130-160): the neighbour's side
// facing back along an edge of type t.
__device__ inline int back_side(uint32_t t) { return (t == ET_FF || t == ET_RF) ? 1 : 0; }

// Initial walk records of a node end (ranking by walks, below): not yet
// walked, or the end of a node that merges with nothing.
constexpr uint64_t kEndUnvisited = ~0ull, kEndSingle = ~0ull - 1;
constexpr uint32_t kWalkCap = 4096;

// Neighbour y of k-mer v (reverse complement rv) for mask bit 4t + x:
// FF w[1:]+x, FR rc(w[1:]+x) = comp(x)+rc(w)[:-1], RF rc(x+w[:-1]) =
// rc(w)[1:]+comp(x), RR x+w[:-1].
template <int W>
__device__ inline KV mask_neighbour(KV v, KV rv, int k, uint32_t bit) {
  const uint32_t t = bit >> 2, x = bit & 3;
  return t == ET_FF ? succ_of<W>(v, k, x)
         : t == ET_FR ? pred_of<W>(rv, k, 3 - x)
         : t == ET_RF ? succ_of<W>(rv, k, 3 - x)
                      : pred_of<W>(v, k, x);
}
// A node listing itself among its neighbours (a tandem repeat).
template <int W>
__device__ inline bool self_listed(KV v, KV rv, int k, uint32_t mask) {
  for (uint32_t m = mask; m; m &= m - 1)
    if (keq<W>(mask_neighbour<W>(v, rv, k, (uint32_t)__builtin_ctz(m)), v)) return true;
  return false;
}

// Per node: tandem flag, and for each side of degree 1 its partner
// (node << 1 | partner side) or kEnd.  The two sides' lookups run
// interleaved, so their loads overlap.  With E given (ranking by walks),
// also each node end's walk entry: the next end on its merge path when
// the link is mergeable (neither node a tandem repeat, the partner's side
// of degree 1; the index slot holds the partner's neighbour mask) and the
// record's initial value, and whether the node ends a path (one mergeable
// side).
// With half (odd k, where no k-mer is its own reverse complement, so each
// node end has one partner end), a link is looked up only by the node of
// the smaller key, which writes the partner entries of both its ends when
// the link is mergeable and nothing otherwise: lnk must start at kEnd, it
// then holds mergeable links only (all its readers want), and k_link_ends
// makes the walk entries from it.
template <int W>
__global__ void k_links(uint64_t nn, const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi,
                        const uint32_t* __restrict__ msk, NodeIndex ix, int k, bool half,
                        uint32_t* __restrict__ lnk, uint8_t* __restrict__ tand, unsigned long long* err,
                        uint4* __restrict__ E, uint32_t* __restrict__ endpoint) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    const KV v = node_key<W>(klo, khi, i);
    const KV rv = revcomp<W>(v, k);
    const uint32_t mask = msk[i] & 0xFFFFu;
    bool tnd = false;
    KV z[2] = {v, v};
    bool need[2] = {false, false};
    uint32_t back[2] = {0, 0};
    for (uint32_t m = mask; m; m &= m - 1) {
      const uint32_t bit = (uint32_t)__builtin_ctz(m);
      const KV y = mask_neighbour<W>(v, rv, k, bit);
      if (keq<W>(y, v)) tnd = true;
      const uint32_t t = bit >> 2;
      const int d = (t == ET_FF || t == ET_FR) ? 0 : 1;
      if (side_deg(mask, d) == 1) {
        z[d] = y;
        // half: the node of the smaller key looks the link up
        need[d] = !half || (W == 1 ? v.lo < y.lo : (v.hi < y.hi || (v.hi == y.hi && v.lo < y.lo)));
        back[d] = (uint32_t)back_side(t);
      }
    }
    // both lookups: bucket bounds, then the probes side by side
    uint64_t s[2] = {0, 0}, R[2] = {0, 0}, base[2] = {0, 0};
    uint32_t found[2] = {kEnd, kEnd}, pmask[2] = {0, 0};
    bool live[2];
    for (int d = 0; d < 2; d++) {
      live[d] = false;
      if (!need[d]) continue;
      const uint32_t b = node_bucket<W>(z[d], ix.nb, ix.bkb);
      const uint32_t lo = ix.dir[b], hi = ix.dir[b + 1];
      if (lo >= hi) continue;
      R[d] = 2ull * (hi - lo);
      base[d] = 2ull * lo;
      s[d] = slot_home(khash<W>(z[d]), R[d]);
      live[d] = true;
    }
    for (uint64_t probe = 0; live[0] || live[1]; probe++) {
      for (int d = 0; d < 2; d++) {
        if (!live[d]) continue;
        if (probe >= R[d]) {
          live[d] = false;
          continue;
        }
        uint64_t key_lo, key_hi = 0;
        uint32_t id, pm;
        if (W == 1) {
          const Slot1 e = ((const Slot1*)ix.slots)[base[d] + s[d]];
          key_lo = e.key;
          id = e.count;
          pm = e.mask;
        } else {
          const Slot2 e = ((const Slot2*)ix.slots)[base[d] + s[d]];
          key_lo = e.lo;
          key_hi = e.hi;
          id = e.count;
          pm = e.mask;
        }
        if (key_lo == z[d].lo && (W == 1 || key_hi == z[d].hi)) {
          found[d] = id;
          pmask[d] = pm;
          live[d] = false;
        } else if ((W == 1 ? key_lo : key_hi) == kEmpty) {
          live[d] = false;
        } else {
          s[d] = s[d] + 1 == R[d] ? 0 : s[d] + 1;
        }
      }
    }
    uint32_t l[2] = {kEnd, kEnd};
    for (int d = 0; d < 2; d++) {
      if (!need[d]) continue;
      if (found[d] == kEnd) atomicMin(err, (unsigned long long)i);
      else l[d] = (found[d] << 1) | back[d];
    }
    tand[i] = tnd ? 1 : 0;
    if (half) {
      // both sides on one partner end would give that end two partners
      const bool same = l[0] != kEnd && l[0] == l[1];
      for (int d = 0; d < 2; d++) {
        if (l[d] == kEnd || same || tnd || found[d] == (uint32_t)i || side_deg(pmask[d], (int)back[d]) != 1 ||
            self_listed<W>(z[d], revcomp<W>(z[d], k), k, pmask[d]))
          continue;
        lnk[2 * i + d] = l[d];
        lnk[2 * (uint64_t)found[d] + back[d]] = ((uint32_t)i << 1) | (uint32_t)d;
      }
      continue;
    }
    lnk[2 * i] = l[0];
    lnk[2 * i + 1] = l[1];
    if (E) {
      bool mg[2];
      uint32_t nx[2];
      for (int d = 0; d < 2; d++) {
        mg[d] = l[d] != kEnd && !tnd && found[d] != (uint32_t)i &&
                side_deg(pmask[d], (int)back[d]) == 1 &&
                !self_listed<W>(z[d], revcomp<W>(z[d], k), k, pmask[d]);
        nx[d] = mg[d] ? ((found[d] << 1) | (1u - back[d])) : kEnd;  // leave the partner by its other side
      }
      const uint64_t r = (mg[0] || mg[1]) ? kEndUnvisited : kEndSingle;
      E[2 * i] = make_uint4(nx[0], 0u, (uint32_t)r, (uint32_t)(r >> 32));
      E[2 * i + 1] = make_uint4(nx[1], 0u, (uint32_t)r, (uint32_t)(r >> 32));
      endpoint[i] = mg[0] != mg[1] ? 1u : 0u;
    }
  }
}

// Walk entries from links that hold mergeable partners only (k_links with
// half): the next end leaves the partner by its other side; the record
// starts unvisited on a node with a mergeable side, single otherwise; a
// node with one mergeable side ends a path.
__global__ void k_link_ends(uint64_t nn, const uint32_t* __restrict__ lnk, uint4* __restrict__ E,
                            uint32_t* __restrict__ endpoint) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint2 l = reinterpret_cast<const uint2*>(lnk)[i];
    const bool m0 = l.x != kEnd, m1 = l.y != kEnd;
    const uint32_t n0 = m0 ? (l.x ^ 1u) : kEnd, n1 = m1 ? (l.y ^ 1u) : kEnd;
    const uint64_t r = (m0 || m1) ? kEndUnvisited : kEndSingle;
    E[2 * i] = make_uint4(n0, 0u, (uint32_t)r, (uint32_t)(r >> 32));
    E[2 * i + 1] = make_uint4(n1, 0u, (uint32_t)r, (uint32_t)(r >> 32));
    endpoint[i] = m0 != m1 ? 1u : 0u;
  }
}

__device__ inline bool mergeable(uint64_t i, int d, const uint32_t* lnk, const uint32_t* msk, const uint8_t* tand,
                                 const uint8_t* cut) {
  uint32_t l = lnk[2 * i + d];
  if (l == kEnd || tand[i] || cut[2 * i + d]) return false;
  uint32_t j = l >> 1;
  int sj = (int)(l & 1);
  if (j == i || tand[j]) return false;
  return side_deg(msk[j] & 0xFFFFu, sj) == 1;
}

// Ranking by walks (odd k).  A merge path is a chain of node ends, each
// leaving its node through one side; next[e] is the end that follows e on
// its path (kEnd at the path's end).  Both endpoints of a path walk it
// toward each other; a walker claims each node it reaches with a
// compare-and-swap on the node's record (walker, side it leaves by,
// position from its start) and stops at the first node the other walker
// claimed, so the two walks split the path and every node is read once.
// A pass over the walkers joins the two halves: the path's length (the sum
// of the two claims) and its smallest node, with the side each walker
// leaves that node by.  A pass over the nodes then turns each record into
// the node's unitig representative (the smallest node; bit 31 when the
// node reads backwards in the unitig, which reads its smallest node
// forwards) and position: the values of the pointer-jumping ranking below.
// Chains are short in read graphs (errors end them every few letters: the
// 5 Mbp workload's longest has ~130 nodes), so one dependent read per node
// replaces several full passes over all ends.  A walker longer than the
// cap, cycles (no endpoint) and even k (a palindromic k-mer links both its
// sides to one neighbour) take the pointer-jumping ranking instead.

// One 16-byte entry per node end: .x = the next end on its path (kEnd at
// the path's end), .z/.w = a 64-bit record; node i's record is the one of
// its F end (entry 2i), in the same line as the entry of its R end, so a
// walk step reads one line.  Record = walker << 32 | side << 31 | position,
// or kEndUnvisited / kEndSingle.
__device__ inline uint64_t end_record(const uint4& x) { return ((uint64_t)x.w << 32) | x.z; }

// The walkers' start ends, in node order.
__global__ void k_walk_list(uint64_t nn, const uint32_t* __restrict__ endpoint, const uint64_t* __restrict__ at,
                            const uint4* __restrict__ E, uint32_t* __restrict__ walkers) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x)
    if (endpoint[i]) walkers[at[i]] = (uint32_t)(2 * i + (E[2 * i].x != kEnd ? 0 : 1));
}

// One walk per path endpoint, claiming nodes until the path's end or a
// node the other walker of the path claimed.  claim[w] = nodes claimed <<
// 32 | (the smallest of them is left by its R side) << 31 | that node
// (0x7fffffff when none); met[w] = the record found on the first node
// another walker claimed (~0 when the walk reached the path's end).
__global__ void __launch_bounds__(kBlock) k_walk(uint32_t nwalk, const uint32_t* __restrict__ walkers,
                                                 uint4* __restrict__ E, uint64_t* __restrict__ claim,
                                                 uint64_t* __restrict__ met, uint32_t* __restrict__ capped) {
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwalk; w += gridDim.x * blockDim.x) {
    uint32_t e = walkers[w], idx = 0, mn = 0x7fffffffu, mn_rev = 0;
    uint64_t other = ~0ull;
    for (;;) {
      const uint32_t i = e >> 1;
      const uint32_t x = E[e].x;
      unsigned long long* rec = reinterpret_cast<unsigned long long*>(E + 2 * (uint64_t)i) + 1;
      const unsigned long long mine = ((unsigned long long)w << 32) | ((unsigned long long)(e & 1) << 31) | idx;
      const unsigned long long old = atomicCAS(rec, (unsigned long long)kEndUnvisited, mine);
      if (old != (unsigned long long)kEndUnvisited) {  // the other walker was here first
        other = old;
        break;
      }
      if (i < mn) {
        mn = i;
        mn_rev = e & 1;
      }
      idx++;
      if (x == kEnd) break;
      if (idx >= kWalkCap) {
        *capped = 1u;
        break;
      }
      e = x;
    }
    claim[w] = ((uint64_t)idx << 32) | ((uint64_t)mn_rev << 31) | mn;
    met[w] = other;
  }
}

// Joins the two walks of each path: a walker that met the other writes
// both walkers' summaries, sum = length << 32 | (the path's smallest node
// is left by its R side, walking as this walker does) << 31 | that node.
// When both met each other, both write the same values.
__global__ void k_walk_join(uint32_t nwalk, const uint64_t* __restrict__ claim, const uint64_t* __restrict__ met,
                            uint64_t* __restrict__ sum) {
  for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwalk; w += gridDim.x * blockDim.x) {
    const uint64_t m = met[w];
    if (m == ~0ull || (m >> 32) >= nwalk) continue;  // not another walker's record: no sum, so pointer jumping
    const uint32_t p = (uint32_t)(m >> 32);
    const uint64_t cw = claim[w], cp = claim[p];
    const uint64_t len = (cw >> 32) + (cp >> 32);
    const uint32_t mw = (uint32_t)cw & 0x7fffffffu, mp = (uint32_t)cp & 0x7fffffffu;
    const uint32_t rw = (uint32_t)(cw >> 31) & 1u, rp = (uint32_t)(cp >> 31) & 1u;
    // the other walker crosses every node the opposite way
    const bool own = mw <= mp;
    const uint32_t mn = own ? mw : mp;
    const uint32_t rev_w = own ? rw : 1u - rp, rev_p = own ? 1u - rw : rp;
    sum[w] = (len << 32) | ((uint64_t)rev_w << 31) | mn;
    sum[p] = (len << 32) | ((uint64_t)rev_p << 31) | mn;
  }
}

// Per node from its record: representative, orientation, position; a node
// without a record lies on a cycle or beyond the cap.  The claiming walker
// reads the unitig backwards when it leaves the smallest node by its R side
// (bit 31 of its sum); the node reads backwards when that differs from the
// side it leaves the node by.
__global__ void k_walk_final(uint64_t nn, const uint4* __restrict__ E, uint64_t nwalk, const uint64_t* __restrict__ sum,
                             uint32_t* __restrict__ urep, uint32_t* __restrict__ upos, uint32_t* __restrict__ isrep,
                             uint64_t* __restrict__ ulen_of_rep, uint32_t* __restrict__ unranked) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t r = end_record(E[2 * i]);
    if (r == kEndSingle) {
      urep[i] = (uint32_t)i;
      upos[i] = 0;
      isrep[i] = 1;
      ulen_of_rep[i] = 1;
      continue;
    }
    if (r == kEndUnvisited || (r >> 32) >= nwalk) {
      *unranked = 1u;
      continue;
    }
    const uint64_t s = sum[r >> 32];
    if (s == ~0ull) {  // no join reached this walker: rank again by pointer jumping
      *unranked = 1u;
      continue;
    }
    const uint32_t len = (uint32_t)(s >> 32), mn = (uint32_t)s & 0x7fffffffu, a = (uint32_t)r & 0x7fffffffu;
    const bool against = ((s >> 31) & 1) != 0;  // the unitig reads against this walker
    const bool side = ((r >> 31) & 1) != 0;      // the walker leaves this node by its R side
    urep[i] = mn | (against != side ? 0x80000000u : 0u);
    upos[i] = against ? len - 1 - a : a;
    isrep[i] = mn == (uint32_t)i ? 1u : 0u;
    ulen_of_rep[i] = mn == (uint32_t)i ? len : 0u;
  }
}

// Packed list ranking: one 16-byte record per element (ptr, dist, min node,
// last element), so a hop is one random access; rounds are double-buffered
// over all elements (own records coalesced), and each block reports whether
// any of its elements is still following its chain.

__global__ void k_pj_init(uint64_t nn, const uint32_t* __restrict__ lnk, const uint32_t* __restrict__ msk,
                          const uint8_t* __restrict__ tand, const uint8_t* __restrict__ cut, uint4* __restrict__ R0) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 2 * nn;
       e += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t i = e >> 1;
    int d = (int)(e & 1);
    uint32_t p = kEnd;
    if (mergeable(i, d, lnk, msk, tand, cut)) {
      uint32_t l = lnk[e];
      p = ((l >> 1) << 1) | (1u - (l & 1));  // leave the partner through its other side
    }
    R0[e] = make_uint4(p, p == kEnd ? 0u : 1u, (uint32_t)i, (uint32_t)e);
  }
}

__global__ void __launch_bounds__(kBlock) k_pj_round(uint64_t ne, const uint4* __restrict__ Rc,
                                                     uint4* __restrict__ Rn, uint8_t* __restrict__ bflag,
                                                     unsigned long long* __restrict__ nread) {
  int any = 0;
  uint32_t reads = 0;  // successor records read (ends still following their chain)
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  const u32x4* __restrict__ Rcv = reinterpret_cast<const u32x4*>(Rc);
  u32x4* __restrict__ Rnv = reinterpret_cast<u32x4*>(Rn);
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (uint64_t)gridDim.x * blockDim.x) {
    // own records stream through (nontemporal), successors are random reads
    u32x4 r = __builtin_nontemporal_load(Rcv + e);
    if (r.x != kEnd) {
      const uint4 q = Rc[r.x];
      r = u32x4{q.x, r.y + q.y, min(r.z, q.z), q.w};
      any |= q.x != kEnd;
      reads++;
    }
    __builtin_nontemporal_store(r, Rnv + e);
  }
  for (int o = 32; o > 0; o >>= 1) reads += __shfl_xor(reads, o);
  if ((threadIdx.x & 63) == 0 && reads) atomicAdd(nread, (unsigned long long)reads);
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) bflag[blockIdx.x] = any ? 1 : 0;
}

__global__ void __launch_bounds__(kBlock) k_flag_or(const uint8_t* __restrict__ bflag, uint32_t n, uint32_t* out) {
  int any = 0;
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) any |= bflag[j];
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) *out = any ? 1u : 0u;
}

// Cycles: the smallest node of each cycle gives up the link on its R side.
__global__ void k_pj_cut(uint64_t ne, const uint4* __restrict__ R, const uint32_t* __restrict__ lnk,
                         uint8_t* __restrict__ cut) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (uint64_t)gridDim.x * blockDim.x) {
    uint4 r = R[e];
    if (r.x == kEnd || (e & 1) == 0 || r.z != (uint32_t)(e >> 1)) continue;
    cut[e] = 1;
    uint32_t l = lnk[e];
    cut[2 * (uint64_t)(l >> 1) + (l & 1)] = 1;
  }
}

__global__ void k_pj_node_final(uint64_t nn, const uint4* __restrict__ R, uint32_t* __restrict__ urep,
                                uint32_t* __restrict__ upos, uint32_t* __restrict__ isrep,
                                uint64_t* __restrict__ ulen_of_rep) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    uint4 f = R[2 * i], b = R[2 * i + 1];
    uint32_t r = min(f.z, b.z);
    uint32_t m = f.y + b.y + 1;
    bool fwd = f.w == R[2 * (uint64_t)r].w;
    uint32_t pos = m - 1 - (fwd ? f.y : b.y);
    urep[i] = r | (fwd ? 0u : 0x80000000u);
    upos[i] = pos;
    isrep[i] = r == (uint32_t)i ? 1 : 0;
    ulen_of_rep[i] = r == (uint32_t)i ? m : 0;
  }
}

// Coverage (saturated at 2^24, as the reference's float sums) of the nodes
// in unitig order; a prefix sum of it gives each unitig's total.
__global__ void k_order_cov(uint64_t nn, const uint32_t* __restrict__ order, const uint32_t* __restrict__ cnt,
                            uint32_t* __restrict__ c) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nn; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = cnt[order[j] & 0x7fffffffu];
    c[j] = x > (1u << 24) ? (1u << 24) : x;
  }
}

__global__ void k_unitig_cov(uint64_t nu, const uint64_t* __restrict__ ubase, const uint64_t* __restrict__ ulen,
                             const uint64_t* __restrict__ cpre, unsigned long long* __restrict__ ucov) {
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nu; u += (uint64_t)gridDim.x * blockDim.x)
    ucov[u] = cpre[ubase[u] + ulen[u]] - cpre[ubase[u]];
}

// Unitig u = rank of its representative; nodes placed in unitig order.
__global__ void k_place(uint64_t nn, const uint32_t* __restrict__ urep, const uint32_t* __restrict__ upos,
                        const uint64_t* __restrict__ uid_of_rep, const uint64_t* __restrict__ ubase,
                        uint32_t* __restrict__ order) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t r = urep[i] & 0x7fffffffu;
    uint64_t u = uid_of_rep[r];
    order[ubase[u] + upos[i]] = (uint32_t)i | (urep[i] & 0x80000000u);
  }
}

__global__ void k_rep_list(uint64_t nn, const uint32_t* __restrict__ isrep, const uint64_t* __restrict__ uid_of_rep,
                           const uint64_t* __restrict__ ulen_of_rep, uint32_t* __restrict__ urep_node,
                           uint64_t* __restrict__ ulen) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!isrep[i]) continue;
    uint64_t u = uid_of_rep[i];
    urep_node[u] = (uint32_t)i;
    ulen[u] = ulen_of_rep[i];
  }
}

// Letter j (0 = first) of node i's k-mer read forward (fwd) or reverse-
// complemented.
template <int W>
__device__ inline uint32_t oriented_letter(const uint64_t* klo, const uint64_t* khi, uint32_t i, bool fwd, int k,
                                           int j) {
  int q = fwd ? j : k - 1 - j;
  uint32_t x = (uint32_t)((q < 32 ? (klo[i] >> (2 * q)) : (W == 2 ? khi[i] >> (2 * (q - 32)) : 0)) & 3);
  return fwd ? x : 3 - x;
}

// Outer edges of unitig u: the 5' end node's outward side and the 3' end
// node's outward side, re-targeted to neighbour unitigs (their representative
// k-mer) and re-typed by the neighbour's orientation; sorted by type then
// neighbour k-mer, duplicates removed.  Up to 16 entries (for even k a
// palindromic neighbour is listed under two types): (type << 30 | rep).
template <int W>
__global__ void k_unitig_edges(uint64_t nu, const uint64_t* __restrict__ ubase, const uint64_t* __restrict__ ulen,
                               const uint32_t* __restrict__ order, const uint64_t* __restrict__ klo,
                               const uint64_t* __restrict__ khi, const uint32_t* __restrict__ msk,
                               NodeIndex ix, const uint32_t* __restrict__ urep, int k,
                               uint32_t* __restrict__ uedge, uint8_t* __restrict__ uecnt) {
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nu; u += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t ent[kUEdges];
    int ne = 0;
    uint64_t m = ulen[u];
    for (int end = 0; end < 2; end++) {
      uint32_t o = order[ubase[u] + (end == 0 ? 0 : m - 1)];
      uint32_t i = o & 0x7fffffffu;
      bool fwd = !(o & 0x80000000u);
      uint32_t uletter = end == 0 ? 1 : 0;              // 5' end -> R, 3' end -> F
      int side = ((end == 0) == fwd) ? 1 : 0;           // node side facing out
      KV v = node_key<W>(klo, khi, i);
      uint32_t mask = msk[i];
      for (uint32_t t = 0; t < 4; t++) {
        int ts = (t == ET_FF || t == ET_FR) ? 0 : 1;
        if (ts != side) continue;
        for (uint32_t x = 0; x < 4; x++) {
          if (!((mask >> (4 * t + x)) & 1)) continue;
          KV z = neighbour<W>(v, k, t, x);
          uint32_t zi = node_lookup<W>(ix, z);
          if (zi == kEnd) continue;  // cannot happen: the graph is symmetric
          uint32_t zr = urep[zi];
          bool z5 = (t == ET_FF || t == ET_RF);           // touches z's 5' end
          bool zfwd = !(zr & 0x80000000u);
          uint32_t wletter = (z5 == zfwd) ? 0 : 1;
          uint32_t ut = uletter == 0 ? (wletter == 0 ? ET_FF : ET_FR) : (wletter == 0 ? ET_RF : ET_RR);
          uint32_t rep = zr & 0x7fffffffu;
          bool dup = false;
          for (int a = 0; a < ne; a++) dup |= (ent[a] >> kRepBits) == ut && (ent[a] & kRepMask) == rep;
          if (!dup && ne < kUEdges) {
            // insertion: by type, then by neighbour k-mer (VKmer byte order,
            // not the node order); the keys are read again rather than kept
            // in a per-thread array, which would live in scratch memory
            const KV rk = node_key<W>(klo, khi, rep);
            int p = ne;
            while (p > 0) {
              uint32_t pt = ent[p - 1] >> kRepBits;
              if (pt < ut || (pt == ut && kle<W>(node_key<W>(klo, khi, ent[p - 1] & kRepMask), rk))) break;
              ent[p] = ent[p - 1];
              p--;
            }
            ent[p] = (ut << kRepBits) | rep;
            ne++;
          }
        }
      }
    }
    for (int a = 0; a < kUEdges; a++) uedge[kUEdges * u + a] = a < ne ? ent[a] : kEnd;
    uecnt[u] = (uint8_t)ne;
  }
}

// Read heads of every node carried into its unitig: offset = k-mer position
// + the head's letter in the k-mer (the build's heads: 0, or k - 1 for
// flipped ones); a reversed node counts the letter from the other end and
// swaps the two sets (Node.mergeUnflippedAndFlippedReadIDs,
// Node.java:740-806).  Sort key:
// hi = unitig, lo = flipped << 53 | offset << 30 | mate << 29 | readId.
__global__ void k_unitig_heads(uint64_t nn, const uint64_t* __restrict__ hstart, const uint32_t* __restrict__ hcu,
                               const uint32_t* __restrict__ hcf, const uint32_t* __restrict__ uval,
                               const uint8_t* __restrict__ hoff, const uint32_t* __restrict__ urep,
                               const uint32_t* __restrict__ upos,
                               const uint64_t* __restrict__ uid_of_rep, const int64_t* __restrict__ rid, int k,
                               Key128* __restrict__ key, uint32_t* __restrict__ val,
                               unsigned long long* __restrict__ too_far) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t nh = hcu[i] + hcf[i];
    if (!nh) continue;
    uint64_t h0 = hstart[i];
    bool fwd = !(urep[i] & 0x80000000u);
    uint64_t u = uid_of_rep[urep[i] & 0x7fffffffu];
    for (uint32_t j = 0; j < nh; j++) {
      uint32_t m = uval[h0 + j] & 0x7fffffffu;
      bool f = j >= hcu[i];
      bool nf = fwd ? f : !f;
      // the head's letter in the node's k-mer: 0 / k - 1 from the build
      const uint32_t a = hoff ? hoff[h0 + j] : (f ? (uint32_t)(k - 1) : 0u);
      uint64_t off = upos[i] + (fwd ? a : (uint32_t)(k - 1) - a);
      if (off > kMaxHeadOffset) {  // ReadHeadInfo.makeUUID throws here (ReadHeadInfo.java:115-119)
        atomicMin(too_far, (unsigned long long)off);
        off = 0;
      }
      Key128 kk;
      kk.hi = u;
      kk.lo = ((uint64_t)nf << 53) | (off << 30) | ((uint64_t)(m & 1) << 29) | (uint64_t)rid[m >> 1];
      key[h0 + j] = kk;
      val[h0 + j] = m | ((uint32_t)nf << 31);
    }
  }
}

__global__ void k_uhead_unique(uint64_t nh, const Key128* __restrict__ key, uint32_t* __restrict__ uflag,
                               uint32_t* __restrict__ hu, uint32_t* __restrict__ hf) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += (uint64_t)gridDim.x * blockDim.x) {
    bool first = j == 0 || key[j].lo != key[j - 1].lo || key[j].hi != key[j - 1].hi;
    uflag[j] = first ? 1 : 0;
    if (first) atomicAdd(((key[j].lo >> 53) & 1) ? &hf[key[j].hi] : &hu[key[j].hi], 1u);
  }
}

__global__ void k_uhead_compact(uint64_t nh, const uint32_t* __restrict__ uflag, const uint64_t* __restrict__ upos_,
                                const Key128* __restrict__ key, const uint32_t* __restrict__ val,
                                uint64_t* __restrict__ ukey_lo, uint32_t* __restrict__ uval_out) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nh; j += (uint64_t)gridDim.x * blockDim.x) {
    if (!uflag[j]) continue;
    ukey_lo[upos_[j]] = key[j].lo;
    uval_out[upos_[j]] = val[j];
  }
}

struct USerCtx {
  int k, nb;
  uint64_t nu;
  const uint64_t *klo, *khi;
  const uint64_t *ubase, *ulen;
  const uint32_t* order;
  const uint32_t* urep_node;
  const unsigned long long* ucov;
  const uint32_t* uedge;
  const uint8_t* uecnt;
  const uint64_t* hstart;  // first head of each unitig
  const uint32_t *hu, *hf;
  const uint64_t* hkey;    // offset etc.
  const uint32_t* hval;    // mate index | flipped << 31
  const uint64_t* hbytes;  // head-rich unitigs: bytes of the unflipped / flipped heads
  const uint32_t* mlen;
  const uint64_t* wbase;
  const uint64_t* words;
  const int64_t* rid;
  const uint8_t* ends;  // per node: letter k-1 | complement of letter 0 << 2
};

// The outer letters of every node, one byte per node (k_node_ends): a
// unitig sequence takes one letter from each node after its first, and a
// gather of these bytes touches far fewer lines than one of the keys.
template <int W>
__global__ void k_node_ends(uint64_t nn, const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi, int k,
                            uint8_t* __restrict__ ends) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    const int q = k - 1;
    const uint32_t last = (uint32_t)((q < 32 ? klo[i] >> (2 * q) : (W == 2 ? khi[i] >> (2 * (q - 32)) : 0)) & 3);
    const uint32_t first = (uint32_t)(klo[i] & 3);
    ends[i] = (uint8_t)(last | ((3 - first) << 2));
  }
}

// Letter k-1 of node o & 0x7fffffff read in the orientation o's top bit
// gives (set: reverse complement), from the node's outer-letter byte.
__device__ inline uint32_t next_letter(const USerCtx& c, uint32_t o) {
  const uint32_t e = c.ends[o & 0x7fffffffu];
  return (o & 0x80000000u) ? e >> 2 : e & 3u;
}

__device__ inline uint64_t useq_len(const USerCtx& c, uint64_t u) { return c.ulen[u] + c.k - 1; }

// Unitigs with more than kHeadsLong read heads have their heads written by
// k_uheads_long (a wave per unitig, lanes over heads); k_useq_write leaves
// the bytes to it.
constexpr uint32_t kHeadsLong = 32;

__device__ inline uint64_t uhead_bytes(const USerCtx& c, uint64_t j) {
  const uint32_t m = c.hval[j] & 0x7fffffffu;
  const uint32_t L = c.mlen[m], Lm = c.mlen[m ^ 1];
  return 1 + 8 + 4 + (L + 3) / 4 + (Lm > 0 ? 4 + (Lm + 3) / 4 : 0);
}

// One read head of unitig records: mate flag, uuid (offset in the unitig,
// mate, readId), the read and its mate as VKmers.
__device__ inline void write_uhead(const USerCtx& c, uint64_t j, uint8_t*& p) {
  const uint32_t mi = c.hval[j] & 0x7fffffffu;
  const uint32_t L = c.mlen[mi], Lm = c.mlen[mi ^ 1];
  *p++ = Lm > 0 ? 1 : 0;
  const uint64_t off = (c.hkey[j] >> 30) & ((1ull << 23) - 1);
  const uint64_t uuid = (off << 40) | ((uint64_t)(mi & 1) << 35) | (uint64_t)c.rid[mi >> 1];
  put64(p, uuid);
  put_read(p, c.words + c.wbase[mi], L);
  if (Lm > 0) put_read(p, c.words + c.wbase[mi ^ 1], Lm);
}

__device__ inline uint64_t unitig_record_size(const USerCtx& c, uint64_t u) {
  uint64_t val = 1 + 4;  // flags + coverage
  int ne = c.uecnt[u];
  int per[4] = {0, 0, 0, 0};
  for (int a = 0; a < ne; a++) per[c.uedge[kUEdges * u + a] >> kRepBits]++;
  for (int t = 0; t < 4; t++)
    if (per[t]) val += 4 + (uint64_t)per[t] * (4 + c.nb);
  uint64_t h0 = c.hstart[u];
  uint32_t nu_ = c.hu[u], nf = c.hf[u];
  for (int f = 0; f < 2; f++) {
    uint32_t n = f ? nf : nu_;
    if (!n) continue;
    val += 1 + 4;
    if (nu_ + nf > kHeadsLong) {  // summed by k_uhead_bytes_long
      val += c.hbytes[2 * u + f];
      continue;
    }
    uint64_t a = h0 + (f ? nu_ : 0);
    for (uint64_t j = a; j < a + n; j++) val += uhead_bytes(c, j);
  }
  if (c.ulen[u] > 1) val += 4 + (useq_len(c, u) + 3) / 4;
  return 8 + 4 + c.nb + val;
}

// Bytes of the heads of head-rich unitigs (more than kHeadsLong), a wave
// per unitig.
__global__ void __launch_bounds__(kBlock) k_uhead_bytes_long(USerCtx c, uint64_t* __restrict__ hbytes) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t u = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); u < c.nu; u += nw) {
    const uint32_t nu_ = c.hu[u], nf = c.hf[u];
    if (nu_ + nf <= kHeadsLong) continue;
    for (int f = 0; f < 2; f++) {
      const uint32_t n = f ? nf : nu_;
      const uint64_t a = c.hstart[u] + (f ? nu_ : 0);
      uint64_t sum = 0;
      for (uint32_t q = lane; q < n; q += 64) sum += uhead_bytes(c, a + q);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
      if (lane == 0) hbytes[2 * u + f] = sum;
    }
  }
}

__global__ void k_useq_size(USerCtx c, uint64_t* __restrict__ rsize) {
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < c.nu; u += (uint64_t)gridDim.x * blockDim.x)
    rsize[u] = unitig_record_size(c, u);
}

// Sequences longer than kUseqLong bases are written by k_useq_long (a wave
// per unitig, lanes over bytes); k_useq_write leaves their bytes to it and
// records where they start.
constexpr uint64_t kUseqLong = 256;
constexpr int kUSerStage = 30720;  // LDS stage of 256 unitig records (longer than node records)

template <int W>
__device__ inline uint8_t useq_byte(const USerCtx& c, const uint32_t* ord, uint64_t len, uint64_t q) {
  // byte q from the least significant end holds bases 4q..4q+3 (letter 0 low)
  uint32_t byte = 0;
  uint32_t o0 = ord[0];
  for (int s2 = 0; s2 < 4; s2++) {
    uint64_t j = 4 * q + s2;
    if (j >= len) break;
    uint32_t x;
    if (j < (uint64_t)c.k) {
      x = oriented_letter<W>(c.klo, c.khi, o0 & 0x7fffffffu, !(o0 & 0x80000000u), c.k, (int)j);
    } else {
      x = next_letter(c, ord[j - c.k + 1]);
    }
    byte |= x << (2 * s2);
  }
  return (uint8_t)byte;
}

// Records of kBlock consecutive unitigs are contiguous: each thread writes
// its record into an LDS stage and the block stores the span with
// coalesced 4-byte writes (spans larger than the stage go direct).
template <int W>
__device__ inline void write_unitig_record(const USerCtx& c, uint64_t u, uint8_t* rs, uint64_t rec,
                                           uint64_t rec_at, uint64_t* __restrict__ seq_at,
                                           uint64_t* __restrict__ head_at) {
  {
    uint8_t* p = rs;
    // the unitig's fields, its edge entries and the first edges' keys are
    // loaded before the first store (stores through p could alias them and
    // would order every later load after them)
    const uint32_t rep = c.urep_node[u];
    const int ne = c.uecnt[u];
    const uint32_t nu_ = c.hu[u], nf = c.hf[u];
    const uint64_t m = c.ulen[u];
    const uint64_t h0 = (nu_ | nf) ? c.hstart[u] : 0ull;
    const unsigned long long ucov = c.ucov[u];
    const uint32_t* ord = c.order + (m > 1 ? c.ubase[u] : 0ull);  // unitig order of its nodes
    const uint32_t o0 = m > 1 ? ord[0] : 0u;                      // first node (with orientation)
    uint32_t ed[kUEdges];
#pragma unroll
    for (int a = 0; a < kUEdges; a++) ed[a] = a < ne ? c.uedge[kUEdges * u + a] : 0u;
    const KV v = node_key<W>(c.klo, c.khi, rep);
    constexpr int kPre = 4;  // edges whose keys are gathered up front
    KV ek[kPre];
#pragma unroll
    for (int a = 0; a < kPre; a++) ek[a] = a < ne ? node_key<W>(c.klo, c.khi, ed[a] & kRepMask) : KV{0, 0};
    put32(p, (uint32_t)(rec - 8));
    put32(p, 4 + c.nb);
    put_key<W>(p, v, c.k, c.nb);
    int per[4] = {0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < kUEdges; a++)
      if (a < ne) per[ed[a] >> kRepBits]++;
    uint8_t flags = 0x80;
    for (int t = 0; t < 4; t++)
      if (per[t]) flags |= 1 << t;
    if (nu_) flags |= 1 << 4;
    if (nf) flags |= 1 << 5;
    if (m > 1) flags |= 1 << 6;
    *p++ = flags;
    // the entries are sorted by type (k_unitig_edges), so each type's list
    // is a run of them: its count before its first entry
    int prev_t = -1;
#pragma unroll
    for (int a = 0; a < kUEdges; a++) {
      if (a < ne) {
        const int t = (int)(ed[a] >> kRepBits);
        if (t != prev_t) {
          put32(p, (uint32_t)per[t]);
          prev_t = t;
        }
        put_key<W>(p, a < kPre ? ek[a] : node_key<W>(c.klo, c.khi, ed[a] & kRepMask), c.k, c.nb);
      }
    }
    for (int f = 0; f < 2; f++) {
      uint32_t n = f ? nf : nu_;
      if (!n) continue;
      *p++ = 1;
      put32(p, n);
      uint64_t a = h0 + (f ? nu_ : 0);
      if (nu_ + nf > kHeadsLong) {  // k_uheads_long writes these heads
        head_at[2 * u + f] = rec_at + (uint64_t)(p - rs);
        p += c.hbytes[2 * u + f];
      } else {
        for (uint64_t j = a; j < a + n; j++) write_uhead(c, j, p);
      }
    }
    if (m > 1) {
      // internal k-mer: the unitig sequence (overlap-(k-1) concatenation,
      // VKmer.mergeWith*Kmer, VKmer.java:525-692) as VKmer bytes
      const uint64_t len = m + c.k - 1;  // useq_len
      put32(p, (uint32_t)len);
      uint64_t nbl = (len + 3) / 4;
      if (len > kUseqLong) {
        seq_at[u] = rec_at + (uint64_t)(p - rs);
        p += nbl;
      } else {
        // bytes from the most significant end, eight per store; the first
        // node's k letters come from its key in registers, each later
        // letter is the last letter of the next node
        const KV v0 = node_key<W>(c.klo, c.khi, o0 & 0x7fffffffu);
        const bool f0 = !(o0 & 0x80000000u);
        uint64_t acc = 0;
        int nacc = 0;
        for (int64_t q = (int64_t)nbl - 1; q >= 0; q--) {
          uint32_t byte = 0;
          for (int s2 = 0; s2 < 4; s2++) {
            const uint64_t j = 4 * (uint64_t)q + s2;
            if (j >= len) break;
            uint32_t x;
            if (j < (uint64_t)c.k) {
              const int qq = f0 ? (int)j : c.k - 1 - (int)j;
              x = (uint32_t)((qq < 32 ? v0.lo >> (2 * qq) : v0.hi >> (2 * (qq - 32))) & 3);
              if (!f0) x = 3 - x;
            } else {
              x = next_letter(c, ord[j - c.k + 1]);
            }
            byte |= x << (2 * s2);
          }
          acc |= (uint64_t)byte << (8 * nacc);
          if (++nacc == 8) {
            memcpy(p, &acc, 8);
            p += 8;
            acc = 0;
            nacc = 0;
          }
        }
        put_le(p, acc, nacc);
      }
    }
    double cov = (double)ucov / (double)m;
    float fc = (float)cov;
    uint32_t bits;
    memcpy(&bits, &fc, 4);
    put32(p, bits);
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_useq_write(USerCtx c, const uint64_t* __restrict__ roff,
                                                       uint8_t* __restrict__ out, uint64_t* __restrict__ seq_at,
                                                       uint64_t* __restrict__ head_at) {
  __shared__ __attribute__((aligned(16))) uint8_t stage[kUSerStage + 8];
  const uint64_t first = (uint64_t)blockIdx.x * kBlock;
  if (first >= c.nu) return;
  const uint64_t last = first + kBlock < c.nu ? first + kBlock : c.nu;
  const uint64_t s0 = roff[first], s1 = roff[last];
  const uint64_t base = s0 & ~3ull;
  const bool staged = (s1 - base) <= kUSerStage;
  const uint64_t u = first + threadIdx.x;
  if (u < last)
    write_unitig_record<W>(c, u, staged ? stage + (roff[u] - base) : out + roff[u], roff[u + 1] - roff[u], roff[u],
                           seq_at, head_at);
  if (!staged) return;
  __syncthreads();
  uint64_t a4 = (s0 + 3) & ~3ull, e4 = s1 & ~3ull;
  if (a4 > e4) { a4 = s1; e4 = s1; }
  for (uint64_t x = s0 + threadIdx.x; x < a4; x += kBlock) out[x] = stage[x - base];
  for (uint64_t x = e4 + threadIdx.x; x < s1; x += kBlock) out[x] = stage[x - base];
  const uint32_t* src = reinterpret_cast<const uint32_t*>(stage + (a4 - base));
  uint32_t* dst = reinterpret_cast<uint32_t*>(out + a4);
  for (uint64_t x = threadIdx.x; x < (e4 - a4) / 4; x += kBlock)
    __builtin_nontemporal_store(src[x], dst + x);  // written once, read by the caller later
}

// Heads of unitigs with many of them: a wave per unitig; each pass places
// 64 heads by a wave prefix sum of their sizes.
__global__ void __launch_bounds__(kBlock) k_uheads_long(USerCtx c, const uint64_t* __restrict__ head_at,
                                                       uint8_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t u = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); u < c.nu; u += nw) {
    const uint32_t nu_ = c.hu[u], nf = c.hf[u];
    if (nu_ + nf <= kHeadsLong) continue;
    for (int f = 0; f < 2; f++) {
      const uint32_t n = f ? nf : nu_;
      if (!n) continue;
      const uint64_t a = c.hstart[u] + (f ? nu_ : 0);
      uint64_t at = head_at[2 * u + f];
      for (uint32_t q0 = 0; q0 < n; q0 += 64) {
        const uint32_t q = q0 + lane;
        const uint32_t sz = q < n ? (uint32_t)uhead_bytes(c, a + q) : 0u;  // < 2^16 per head
        uint32_t inc = sz;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(inc, o, 64);
          if (lane >= o) inc += y;
        }
        if (q < n) {
          uint8_t* p = out + at + (inc - sz);
          write_uhead(c, a + q, p);
        }
        at += (uint32_t)__shfl(inc, 63, 64);
      }
    }
  }
}

template <int W>
__global__ void __launch_bounds__(kBlock) k_useq_long(USerCtx c, const uint64_t* __restrict__ seq_at,
                                                     uint8_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t u = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); u < c.nu; u += nw) {
    if (c.ulen[u] <= 1) continue;
    const uint64_t len = useq_len(c, u);
    if (len <= kUseqLong) continue;
    const uint64_t nbl = (len + 3) / 4;
    const uint32_t* ord = c.order + c.ubase[u];
    uint8_t* p = out + seq_at[u];
    for (uint64_t bq = lane; bq < nbl; bq += 64) p[bq] = useq_byte<W>(c, ord, len, nbl - 1 - bq);
  }
}

}  // namespace gx

// ===========================================================================
// Host side
// ===========================================================================

using namespace gx;

constexpr int kMaxTimedRounds = 64;  // pointer-jumping rounds timed per unitig run

struct gx_build {
  gx_build_cfg cfg{};
  int W = 1, nb = 0;
  hipStream_t st = nullptr;
  // a second stream for work that overlaps the main stream's (the replay
  // of the longest record runs beside the heads, the unitigs' outer edges
  // beside their coverage and heads): ordered by events, created on first
  // use
  hipStream_t st2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;
  uint64_t text_len = 0;
  int fmt = -1;                     // input format fixed by the first add call (GX_FMT_*)
  DevBuf text2, rtext, fterm1, fterm2, fsz, foff;  // second FASTQ file; readids text made from FASTQ
  uint64_t text2_len = 0, rtext_len = 0;
  const uint8_t* src = nullptr;     // readids text the build parses
  uint64_t src_len = 0;
  bool finished = false;
  // device buffers (grown on demand, reused across builds)
  DevBuf text, scratch, blk, blk2, term, rid, mstart, mlen, wcnt, wbase, words, valid, nocc, occbase, hmask, hidx;
  DevBuf table, occ_slot, occ_code, claimed;
  DevBuf ckey, cval, ckey2, cval2, nslot, klo, khi, cnt, msk, needs, selcnt, seg;
  DevBuf selflag, selpos, selk, selv, selk2, selv2, lists, rpos, rkey, rval, rkey2, rval2;
  DevBuf hkey, hval, hkey2, hval2, uflag, upos, uval, hcu, hcf, htot, hstart, rsize, roff, out;
  DevBuf errw;
  uint64_t stats[16] = {};
  // per-phase device timings (hipEvents on the handle's stream)
  bool prof = false;
  hipEvent_t evpj[2 * kMaxTimedRounds] = {};  // per pointer-jumping round
  int npj = 0;
  hipEvent_t ev[2 * 16] = {};
  bool ev_used[16] = {};
  double phase_ms[16] = {};
  // unitig pass
  DevBuf u_lnk, u_tand, u_cut, u_ptr[2], u_flag, u_any, u_rep, u_pos, u_isrep, u_lenrep, u_walk, u_walk2;
  DevBuf u_uid, u_repnode, u_len, u_base, u_order, u_cov, u_edge, u_ecnt, u_hk, u_hk2, u_hv, u_hv2, u_uflag, u_uposh;
  DevBuf u_hu, u_hf, u_htot, u_hstart, u_hklo, u_hvc, u_rsize, u_roff, u_out;
  uint64_t nuheads = 0, u_bytes = 0, u_count = 0, u_rounds = 0, u_succ_reads = 0, u_walkers = 0;
  bool no_walk = false;  // rank by pointer jumping only (GX_UNITIG_RANK_POINTER_JUMPING)
  bool unitigs_done = false;
  uint64_t out_bytes = 0, nnodes = 0;
  DevBuf ppid, pidx, ppid2, pidx2, pbound;
  DevBuf sva, svb;  // sorted aggregation: record neighbour codes
  DevBuf nix_dir, nix_slots;  // key -> node index
  DevBuf rcode;               // replay codes, one byte per record in node order
  DevBuf u_seqat;             // unitig pass: byte position of long sequences
  DevBuf u_hbytes;            // unitig pass: head bytes of head-rich unitigs
  DevBuf u_headat;            // unitig pass: byte position of the heads of head-rich unitigs
  DevBuf u_ends;              // unitig pass: each node's outer letters, one byte per node
  DevBuf s_acc, s_k1, s_k2, s_v1, s_v2, s_flag, s_pos, s_bval, s_bst, s_out;  // statistics, FASTA
  int st_graph = -1;                                                          // graph of the last statistics
  std::vector<uint64_t> st_tot;
  std::vector<int64_t> st_sum, st_max;
  std::vector<std::vector<std::pair<int64_t, uint64_t>>> st_bins;
  DevBuf tophist, outacc;     // key ranges: top-byte histogram; records of earlier ranges
  DevBuf blist;               // record sort: buckets listed for the larger LDS tiers
  DevBuf hist2;               // key ranges: the next range's per-tile digit counts
  // graph cleaning: the compacted graph; key ranges: the node graph of the ranges so far
  DevBuf c_keep, c_pos, c_klo, c_khi, c_cnt, c_msk, c_hcu, c_hcf, c_htot, c_hstart, c_uval;
  DevBuf hoff, c_hoff;  // per node head: its letter in the node's k-mer (after a bubble merge)
  DevBuf bb_min, bb_maj, bb_typ, bb_np, bb_pbase, bb_pairs, bb_ed, bb_into, bb_hn, bb_hb, bb_addu, bb_addf, bb_cur,
      bb_tot, bb_astart, bb_hnode, bb_hmv, bb_hofs, bb_aval, bb_aoff;  // bubble merge
  uint64_t bub_pairs = 0;
  bool has_hoff = false;
  uint64_t range_occ = 0;     // occurrences per key-range pass (0: from the HBM size)
  gx_part_sink sink = nullptr;
  void* sink_user = nullptr;
  bool streamed = false;      // the last build's records went to the sink
  uint32_t nranges = 1;
  bool graph_held = true;     // the node graph of the last build is in HBM (not after sink passes)
  int nix_bkb = 2;            // key bytes of the node index directory
  ReclaimList reclaim;        // buffers the last build reads no more
  // learned sizing
  gx_mlp* sizer = nullptr;
  unsigned grid = 2048;  // blocks of grid-stride kernels (8 per CU)
  float slack = 1.5f;    // learned table capacity / predicted distinct
  int method = GX_AGG_SORT;
  DevBuf hregs, sfeat, spred;
  float feat[GX_SIZER_FEATURES] = {};
  uint8_t regs_h[GX_SIZER_HLL_REGISTERS] = {};
  double pred[2] = {0, 0};
  std::vector<uint64_t> part_rec, part_off;  // per partition: first record, first byte (P + 1 entries)
};

static thread_local std::string g_create_err;  // errors of calls without a handle (per thread)

enum Phase {
  PH_PARSE, PH_INSERT, PH_NODES, PH_REPLAY, PH_HEADS, PH_SER, PH_KINSERT, PH_KREPLAY, PH_UNITIG, PH_SIZING, PH_NIX,
  PH_KPJ, PH_KSCATTER, PH_KWALK, NPH
};
static const char* kPhaseNames[NPH] = {"parse", "insert", "sort_nodes", "replay", "heads", "serialize", "k_insert",
                                       "k_replay", "unitig", "sizing", "node_index", "k_pj_round", "k_occ_scatter",
                                       "k_walk"};

// s: the stream the phase runs on (default: the main stream)
static void ph_begin(gx_build* b, int ph, hipStream_t s = nullptr) {
  if (!b->prof) return;
  if (!b->ev[2 * ph]) {
    GX_HIP(hipEventCreate(&b->ev[2 * ph]));
    GX_HIP(hipEventCreate(&b->ev[2 * ph + 1]));
  }
  GX_HIP(hipEventRecord(b->ev[2 * ph], s ? s : b->st));
  b->ev_used[ph] = true;
}
static void ph_end(gx_build* b, int ph, hipStream_t s = nullptr) {
  if (b->prof) GX_HIP(hipEventRecord(b->ev[2 * ph + 1], s ? s : b->st));
}
// add: sum into the phase totals (a build of several key ranges)
static void ph_collect(gx_build* b, bool add = false) {
  for (int i = 0; i < NPH; i++) {
    if (b->prof && b->ev_used[i]) {
      float ms = 0;
      GX_HIP(hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]));
      b->phase_ms[i] = add ? b->phase_ms[i] + ms : ms;
    }
    b->ev_used[i] = false;
  }
  if (b->prof && b->npj) {
    double tot = 0;
    for (int r = 0; r < b->npj; r++) {
      float ms = 0;
      GX_HIP(hipEventElapsedTime(&ms, b->evpj[2 * r], b->evpj[2 * r + 1]));
      tot += ms;
    }
    b->phase_ms[PH_KPJ] = tot;
  }
  b->npj = 0;
}

// events around each launch of a repeated kernel (summed at collect)
static void pj_begin(gx_build* b) {
  if (!b->prof || b->npj >= kMaxTimedRounds) return;
  int r = b->npj;
  if (!b->evpj[2 * r]) {
    GX_HIP(hipEventCreate(&b->evpj[2 * r]));
    GX_HIP(hipEventCreate(&b->evpj[2 * r + 1]));
  }
  GX_HIP(hipEventRecord(b->evpj[2 * r], b->st));
}
static void pj_end(gx_build* b) {
  if (!b->prof || b->npj >= kMaxTimedRounds) return;
  GX_HIP(hipEventRecord(b->evpj[2 * b->npj + 1], b->st));
  b->npj++;
}

static char* scratch_for(gx_build* b, size_t bytes) { return b->scratch.ensure(bytes) ? b->scratch.as<char>() : nullptr; }

template <class T>
static T* buf(DevBuf& d, size_t n) {
  return static_cast<T*>(d.ensure(n * sizeof(T)));
}

static void side_init(gx_build* b) {
  if (b->st2) return;
  GX_HIP(hipStreamCreateWithFlags(&b->st2, hipStreamNonBlocking));
  b->reclaim.st2 = b->st2;
  for (hipEvent_t* e : {&b->ev_fork, &b->ev_join})
    GX_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
}
// Work enqueued on waiter after this call starts after everything enqueued
// on producer so far.
static void stream_after(hipStream_t waiter, hipStream_t producer, hipEvent_t ev) {
  GX_HIP(hipEventRecord(ev, producer));
  GX_HIP(hipStreamWaitEvent(waiter, ev, 0));
}

template <class T>
static T read1(hipStream_t st, const T* dptr) {
  T v;
  GX_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, st));
  GX_HIP(hipStreamSynchronize(st));
  return v;
}

static const char* errname(uint32_t code) {
  switch (code) {
    case ERR_FIELDS:
      return "input format is not correct! only support id'\\t'readSeq'\\t'mateReadSeq or id'\\t'readSeq";
    case ERR_READID: return "read id is not a valid long";
    case ERR_READID_RANGE: return "readId will lose some of its bits when saved (ReadHeadInfo allows 29 bits)";
    case ERR_SHORT: return "kmersize is larger than the read length";
    case ERR_TABLE: return "hash table overflow";
  }
  return "input error";
}

static void check_err(gx_build* b) {
  unsigned long long e = read1(b->st, b->errw.as<unsigned long long>());
  if (e != ~0ull) {
    uint32_t code = (uint32_t)(e & 0xff);
    uint64_t line = e >> 8;
    if (code == ERR_TABLE) throw GxError(GX_ENOMEM, "hash table overflow");
    throw GxError(GX_EINPUT, std::string(errname(code)) + " (line " + std::to_string(line + 1) + ")");
  }
}

// Sizing features of a read set from the HLL registers of its first mates
// (the network's input layout):
//   [0, 64)  histogram of register values / 4096, shifted so that the
//            rounded mean register value lands in bin 32 (scale-free);
//   64       HLL estimate of the sample's distinct k-mers / its occurrences;
//   65       log1p(coverage estimate) / 5, coverage estimate =
//            2 * max(0, 1 - estimate / occurrences) / sample fraction;
//   66       k / 64;  67  mean valid mate length / 256;
//   68       sample fraction * 10;  69  log2(partitions) / 16;
//   70       0 (reserved);  71  1.
// HLL estimate: alpha_m m^2 / sum 2^-M[j], linear counting m ln(m / zeros)
// below 2.5 m (Flajolet et al. 2007).
static void sizing_features(const uint8_t* regs, double occ_s, double valid_s, double frac, int k, int P,
                            float* f) {
  const int m = GX_SIZER_HLL_REGISTERS;
  double hist[64] = {0}, inv = 0, sum = 0;
  int zeros = 0;
  for (int j = 0; j < m; j++) {
    int r = regs[j] < 63 ? regs[j] : 63;
    hist[r] += 1;
    inv += std::ldexp(1.0, -(int)regs[j]);
    sum += regs[j];
    zeros += regs[j] == 0;
  }
  double E = 0.7213 / (1 + 1.079 / m) * m * (double)m / inv;
  if (E <= 2.5 * m && zeros > 0) E = m * std::log((double)m / zeros);
  int center = (int)std::lround(sum / m);
  for (int i = 0; i < GX_SIZER_FEATURES; i++) f[i] = 0.f;
  for (int v = 0; v < 64; v++) {
    int bin = v - center + 32;
    bin = bin < 0 ? 0 : bin > 63 ? 63 : bin;
    f[bin] += (float)(hist[v] / m);
  }
  double ratio = occ_s > 0 ? std::min(2.0, E / occ_s) : 1.0;
  double cov = frac > 0 ? 2.0 * std::max(0.0, 1.0 - ratio) / frac : 0.0;
  f[64] = (float)ratio;
  f[65] = (float)(std::log1p(cov) / 5.0);
  f[66] = (float)(k / 64.0);
  f[67] = (float)((valid_s > 0 ? occ_s / valid_s + k - 1 : 0.0) / 256.0);
  f[68] = (float)(frac * 10.0);
  f[69] = (float)(std::log2((double)P) / 16.0);
  f[70] = 0.f;
  f[71] = 1.f;
}

// Appends len bytes at dst + dlen, growing dst (to at least `hint` bytes
// when that is larger) with its contents kept.
static void append_to(gx_build* b, DevBuf& dst, uint64_t& dlen, const void* src, size_t len, hipMemcpyKind kind,
                      size_t hint = 0);
static void swap_buf(DevBuf& x, DevBuf& y) {
  std::swap(x.p, y.p);
  std::swap(x.cap, y.cap);
}

// Key -> node index over nodes [0, nn) in key order whose keys' top byte
// lies in [tlo, thi).
template <int W>
static NodeIndex build_node_index(gx_build* b, hipStream_t st, uint64_t nn, const uint64_t* klo, const uint64_t* khi,
                                  const uint32_t* msk, uint32_t tlo, uint32_t thi) {
  const int nb = b->nb;
  const unsigned G = b->grid;
  const size_t slotsz = W == 1 ? sizeof(Slot1) : sizeof(Slot2);
  const int xbkb = nb < 2 ? 1 : (nb >= 4 && nn > (1ull << 26)) ? 3 : 2;
  const uint32_t nbk = 1u << (8 * xbkb);
  const uint32_t xb0 = tlo << (8 * (xbkb - 1)), xb1 = thi << (8 * (xbkb - 1));
  b->nix_bkb = xbkb;
  uint32_t* ndir = buf<uint32_t>(b->nix_dir, nbk + 1);
  void* nslots = b->nix_slots.ensure(std::max<uint64_t>(1, 2 * nn) * slotsz);
  if (xb0 > 0) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(xb0, kBlock, G)), dim3(kBlock), 0, st, ndir, 0, xb0, 0u);
    GX_HIP(hipGetLastError());
  }
  if (xb1 < nbk) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nbk - xb1, kBlock, G)), dim3(kBlock), 0, st, ndir, (uint64_t)xb1 + 1,
                       (uint64_t)nbk + 1, (uint32_t)nn);
    GX_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_nix_dir<W>, dim3(grid_for(nn + 1, kBlock, G)), dim3(kBlock), 0, st, nn, klo, khi, nb, xbkb, xb0,
                     xb1, ndir);
  GX_HIP(hipGetLastError());
  if (xbkb == 3)
    hipLaunchKernelGGL((k_nix_build<W, kNixLdsSmall>), dim3(std::min(xb1 - xb0, 65536u)), dim3(kBlock), 0, st, klo,
                       khi, msk, ndir, nslots, xb0, xb1);
  else
    hipLaunchKernelGGL((k_nix_build<W, kNixLds>), dim3(std::min(xb1 - xb0, 65536u)), dim3(kBlock), 0, st, klo, khi,
                       msk, ndir, nslots, xb0, xb1);
  GX_HIP(hipGetLastError());
  return NodeIndex{ndir, nslots, nb, nbk, xbkb};
}

// Device bytes the handle holds (all of its buffers).
// Every device buffer of a handle.
template <class F>
static void for_each_buf(gx_build* b, F&& f) {
  f(&b->u_ptr[0]);
  f(&b->u_ptr[1]);
  for (DevBuf* d :
       {&b->text2, &b->rtext, &b->fterm1, &b->fterm2, &b->fsz, &b->foff, &b->text, &b->scratch, &b->blk, &b->blk2,
        &b->term, &b->rid, &b->mstart, &b->mlen, &b->wcnt, &b->wbase, &b->words, &b->valid, &b->nocc, &b->occbase,
        &b->hmask, &b->hidx, &b->table, &b->occ_slot, &b->occ_code, &b->claimed, &b->ckey, &b->cval, &b->ckey2,
        &b->cval2, &b->nslot, &b->klo, &b->khi, &b->cnt, &b->msk, &b->needs, &b->selcnt, &b->seg, &b->selflag,
        &b->selpos, &b->selk, &b->selv, &b->selk2, &b->selv2, &b->lists, &b->rpos, &b->rkey, &b->rval, &b->rkey2,
        &b->rval2, &b->hkey, &b->hval, &b->hkey2, &b->hval2, &b->uflag, &b->upos, &b->uval, &b->hcu, &b->hcf,
        &b->htot, &b->hstart, &b->rsize, &b->roff, &b->out, &b->errw, &b->u_lnk, &b->u_tand, &b->u_cut, &b->u_flag,
        &b->u_any, &b->u_rep, &b->u_pos, &b->u_isrep, &b->u_lenrep, &b->u_uid, &b->u_repnode, &b->u_len,
        &b->u_base, &b->u_order, &b->u_cov, &b->u_edge, &b->u_ecnt, &b->u_hk, &b->u_hk2, &b->u_hv, &b->u_hv2,
        &b->u_uflag, &b->u_uposh, &b->u_hu, &b->u_hf, &b->u_htot, &b->u_hstart, &b->u_hklo, &b->u_hvc, &b->u_rsize,
        &b->u_roff, &b->u_out, &b->ppid, &b->pidx, &b->ppid2, &b->pidx2, &b->pbound, &b->sva, &b->svb,
        &b->nix_dir, &b->nix_slots, &b->rcode, &b->u_seqat, &b->u_hbytes, &b->tophist, &b->outacc, &b->blist, &b->hist2,
        &b->hregs, &b->sfeat, &b->spred, &b->c_keep, &b->c_pos, &b->c_klo, &b->c_khi, &b->c_cnt, &b->c_msk,
        &b->c_hcu, &b->c_hcf, &b->c_htot, &b->c_hstart, &b->c_uval, &b->hoff, &b->c_hoff, &b->bb_min, &b->bb_maj,
        &b->bb_typ, &b->bb_np, &b->bb_pbase, &b->bb_pairs, &b->bb_ed, &b->bb_into, &b->bb_hn, &b->bb_hb,
        &b->bb_addu, &b->bb_addf, &b->bb_cur, &b->bb_tot, &b->bb_astart, &b->bb_hnode, &b->bb_hmv, &b->bb_hofs,
        &b->bb_aval, &b->bb_aoff, &b->u_headat, &b->s_acc, &b->s_k1, &b->s_k2,
        &b->s_v1, &b->s_v2, &b->s_flag, &b->s_pos, &b->s_bval, &b->s_bst, &b->s_out, &b->u_walk, &b->u_walk2,
        &b->u_ends})
    f(d);
}

static size_t handle_bytes(gx_build* b) {
  size_t s = 0;
  for_each_buf(b, [&](DevBuf* d) { s += d->cap; });
  return s;
}

// At the start of a build every buffer but the input is the last build's
// and may be freed when an allocation would otherwise fail; a buffer the
// build uses leaves the list on its first use (DevBuf::ensure).
static void results_dead(gx_build* b) {
  auto& dead = b->reclaim.dead;
  dead.clear();
  for_each_buf(b, [&](DevBuf* d) {
    if (d != &b->text && d != &b->text2 && d != &b->rtext && d != &b->fterm1 && d != &b->fterm2 && d != &b->fsz &&
        d != &b->foff && d->p)
      dead.push_back(d);
  });
}

template <int W>
static void run_build(gx_build* b) {
  hipStream_t st = b->st;
  for (int i = 0; i < NPH; i++)
    if (i != PH_UNITIG && i != PH_KPJ) b->phase_ms[i] = 0;
  const int k = b->cfg.k, nb = b->nb;
  const uint64_t n = b->src_len;
  const uint8_t* t = b->src;
  const unsigned G = b->grid;  // grid for grid-stride kernels
  unsigned long long* errw = buf<unsigned long long>(b->errw, 4);
  GX_HIP(hipMemsetAsync(errw, 0xff, 16, st));
  GX_HIP(hipMemsetAsync(errw + 2, 0, 16, st));

  // 1. lines
  ph_begin(b, PH_PARSE);
  unsigned lblocks = grid_for(n + 1, kLTile, 0xffffffffu);
  uint64_t* blk = buf<uint64_t>(b->blk, (size_t)lblocks + 1);
  uint64_t* blkb = buf<uint64_t>(b->blk2, (size_t)lblocks + 1);
  hipLaunchKernelGGL(k_line_count, dim3(lblocks), dim3(kBlock), 0, st, t, n, blk, errw + 1);
  GX_HIP(hipGetLastError());
  exclusive_scan<uint64_t, uint64_t>(blk, blkb, lblocks, scratch_for(b, scan_scratch_bytes<uint64_t>(lblocks)), st);
  uint64_t nlines = read1(st, blkb + lblocks);
  unsigned long long na = read1(st, errw + 1);
  if (na != ~0ull) throw GxError(GX_EINPUT, "non-ASCII byte at offset " + std::to_string(na) + " (input must be ASCII)");
  uint64_t* term = buf<uint64_t>(b->term, nlines + 1);
  hipLaunchKernelGGL(k_line_write, dim3(lblocks), dim3(kBlock), 0, st, t, n, blkb, term, errw + 1);
  GX_HIP(hipGetLastError());

  // 2. fields
  uint64_t nmates = 2 * nlines;
  int64_t* rid = buf<int64_t>(b->rid, nlines + 1);
  uint64_t* mstart = buf<uint64_t>(b->mstart, nmates + 1);
  uint32_t* mlen = buf<uint32_t>(b->mlen, nmates + 1);
  if (nlines) {
    unsigned g = grid_for(nlines, kBlock / 64, 1u << 16);
    hipLaunchKernelGGL(k_parse_lines, dim3(g), dim3(kBlock), 0, st, t, n, term, nlines, rid, mstart, mlen, errw);
    GX_HIP(hipGetLastError());
  }
  // 3. pack
  uint64_t* wcnt = buf<uint64_t>(b->wcnt, nmates + 1);
  uint64_t* wbase = buf<uint64_t>(b->wbase, nmates + 1);
  if (nmates) {
    hipLaunchKernelGGL(k_word_counts, dim3(grid_for(nmates, kBlock, G)), dim3(kBlock), 0, st, nmates, mlen, wcnt);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(wcnt, wbase, nmates, scratch_for(b, scan_scratch_bytes<uint64_t>(nmates)), st);
  uint64_t nwords = read1(st, wbase + nmates);
  uint64_t* words = buf<uint64_t>(b->words, nwords + 4);
  uint8_t* valid = buf<uint8_t>(b->valid, nmates + 1);
  uint64_t* nocc = buf<uint64_t>(b->nocc, nmates + 1);
  if (nmates) {
    unsigned g = grid_for(nmates, kBlock / 64, 1u << 16);
    hipLaunchKernelGGL(k_pack_mates, dim3(g), dim3(kBlock), 0, st, t, nmates, mstart, mlen, wbase, rid, k, words,
                       valid, nocc, errw);
    GX_HIP(hipGetLastError());
  }
  ph_end(b, PH_PARSE);
  check_err(b);
  uint64_t* occbase = buf<uint64_t>(b->occbase, nmates + 1);
  exclusive_scan<uint64_t, uint64_t>(nocc, occbase, nmates, scratch_for(b, scan_scratch_bytes<uint64_t>(nmates)), st);
  uint64_t N = read1(st, occbase + nmates);
  uint64_t* hm = buf<uint64_t>(b->hmask, nmates + 1);
  uint64_t* hidx = buf<uint64_t>(b->hidx, nmates + 1);
  if (nmates) {
    hipLaunchKernelGGL(k_valid_u64, dim3(grid_for(nmates, kBlock, G)), dim3(kBlock), 0, st, nmates, valid, nocc, hm);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(hm, hidx, nmates, scratch_for(b, scan_scratch_bytes<uint64_t>(nmates)), st);
  uint64_t nheads = read1(st, hidx + nmates);

  // 3b. sizing features: HLL sketch of the first 1% of the mates (at least
  //     512), then, with a network attached, the predicted distinct count
  ph_begin(b, PH_SIZING);
  const uint64_t ns_m = std::min<uint64_t>(nmates, std::max<uint64_t>((nmates + 99) / 100, 512));
  uint32_t* hr = buf<uint32_t>(b->hregs, kHllRegs);
  GX_HIP(hipMemsetAsync(hr, 0, kHllRegs * 4, st));
  if (ns_m) {
    hipLaunchKernelGGL(k_hll<W>, dim3(grid_for(ns_m, kBlock / 64, 1024)), dim3(kBlock), 0, st, ns_m, wbase, words,
                       nocc, k, hr);
    GX_HIP(hipGetLastError());
  }
  const uint64_t occ_s = read1(st, occbase + ns_m), valid_s = read1(st, hidx + ns_m);
  {
    std::vector<uint32_t> rh(kHllRegs);
    GX_HIP(hipMemcpyAsync(rh.data(), hr, kHllRegs * 4, hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    for (int i = 0; i < kHllRegs; i++) b->regs_h[i] = (uint8_t)rh[i];
  }
  sizing_features(b->regs_h, (double)occ_s, (double)valid_s, nmates ? (double)ns_m / (double)nmates : 0.0, k,
                  b->cfg.nparts, b->feat);
  b->pred[0] = b->pred[1] = 0;
  uint64_t learned_cap = 0;
  if (b->sizer && !b->cfg.table_hint && N) {
    float* df = buf<float>(b->sfeat, GX_SIZER_FEATURES);
    float* dy = buf<float>(b->spred, 8);
    GX_HIP(hipMemcpyAsync(df, b->feat, sizeof(b->feat), hipMemcpyHostToDevice, st));
    int rc = gx_mlp_forward_device(b->sizer, df, 1, dy, st);
    if (rc) throw GxError(rc, std::string("sizing network: ") + gx_mlp_last_error(b->sizer));
    float y[2];
    GX_HIP(hipMemcpyAsync(y, dy, sizeof(y), hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    double ly = std::isfinite(y[0]) ? std::min(0.0, std::max(-30.0, (double)y[0])) : 0.0;
    b->pred[0] = std::max(1.0, std::floor((double)N * std::exp(ly)));
    b->pred[1] = std::isfinite(y[1]) ? std::exp(std::min(10.0, std::max(-10.0, (double)y[1]))) : 1.0;
    learned_cap = std::max<uint64_t>(1024, (uint64_t)std::ceil((double)b->slack * b->pred[0]));
  }
  ph_end(b, PH_SIZING);

  // 4-5. aggregation: KmerPointable-ordered nodes (key, count, neighbour
  //      mask) and, for the JDK-7 replay, each node's occurrence codes in
  //      emission order; plus a key -> node table for lookups by key.
  using SK = typename SortKeyOf<W>::T;
  const bool sorted = b->method == GX_AGG_SORT;
  const size_t slotsz = W == 1 ? sizeof(Slot1) : sizeof(Slot2);
  const uint32_t P = (uint32_t)b->cfg.nparts;

  // 4'. key ranges.  The sorted method takes the occurrences in passes over
  //     ranges of the sort key's top byte when one pass would not fit in
  //     HBM (or when the caller bounds the occurrences per pass): every node
  //     lies in one range, the ranges follow KmerPointable order, and each
  //     range's records are serialised and handed on (to the sink) or kept
  //     in HBM with the range's node graph before the next.
  std::vector<uint32_t> tb{0u, 256u};
  std::vector<uint64_t> rocc{N};
  std::vector<uint64_t> toph(256, 0);  // occurrences per top key byte (several ranges)
  const bool keep = !b->sink;          // several ranges keep their records and node graph
  DevBuf* const kept_bufs[8] = {&b->outacc, &b->c_klo, &b->c_khi, &b->c_cnt,
                                &b->c_msk,  &b->c_hcu, &b->c_hcf, &b->c_uval};
  const double tiles = (double)((N + kSortTile - 1) / kSortTile);
  // HBM a pass may use: free now plus what this handle holds beyond the
  // parsed reads (its buffers from an earlier build are reused) and beyond
  // `kept` bytes this build holds for later, less a margin
  auto pass_budget = [&](double kept) {
    size_t fr = 0, tot = 0;
    GX_HIP(hipMemGetInfo(&fr, &tot));
    size_t live = 0;
    for (DevBuf* d : {&b->text, &b->text2, &b->rtext, &b->term, &b->rid, &b->mstart, &b->mlen, &b->wcnt, &b->wbase,
                      &b->words, &b->valid, &b->nocc, &b->occbase, &b->hmask, &b->hidx})
      live += d->cap;
    return (double)fr + (double)handle_bytes(b) - (double)live - kept - 0.1 * (double)tot - 48.0 * (double)nheads -
           16.0 * 256.0 * tiles;
  };
  // as few ranges of n occurrences as `budget` allows, of even size: filling
  // the first ranges up to the budget leaves the largest buffers at the edge
  // of HBM (dead buffers freed and allocated again every build) for no
  // fewer passes
  auto range_target = [](uint64_t n, double budget, double per_occ) {
    uint64_t target = budget > per_occ ? (uint64_t)(budget / per_occ) : 1;
    if (target < n) {
      const uint64_t r_min = (n + target - 1) / target;
      target = std::min<uint64_t>(target, (uint64_t)(1.1 * (double)n / (double)r_min) + 1);
    }
    return target;
  };
  // ranges r, r+1, ... from top byte tb[r] on, each of at most `target`
  // occurrences (a single byte value may hold more)
  auto cut_ranges = [&](size_t r, uint64_t target) {
    tb.resize(r + 1);
    rocc.resize(r);
    uint64_t cur = 0;
    for (uint32_t t = tb[r]; t < 256; t++) {
      if (cur && cur + toph[t] > target) {  // close the range before this byte value
        tb.push_back(t);
        rocc.push_back(cur);
        cur = 0;
      }
      cur += toph[t];
    }
    tb.push_back(256u);
    rocc.push_back(cur);
  };
  if (sorted && N) {
    uint64_t target = b->range_occ;
    if (!target) {
      // distinct k-mers per occurrence: the attached sizing network's
      // prediction, else 0.6 (resequencing depths); ranges after the first
      // are planned again from what the first ones measured
      const double dfrac = b->pred[0] > 0 ? std::min(1.0, std::max(0.05, b->pred[0] / (double)N)) : 0.6;
      const double per_occ = pass_bytes_per_occurrence(W, nb, dfrac);
      const double budget = pass_budget(0);
      target = range_target(N, budget, per_occ);
      if (target < N && keep)  // several ranges: room for what they keep
        target = range_target(N, budget - (double)N * dfrac * kept_bytes_per_node(W, nb), per_occ);
    }
    if (target < N) {
      unsigned long long* th = buf<unsigned long long>(b->tophist, 256);
      GX_HIP(hipMemsetAsync(th, 0, 256 * 8, st));
      const unsigned ntiles = (unsigned)((N + kSortTile - 1) / kSortTile);
      hipLaunchKernelGGL(k_top_hist<W>, dim3(ntiles), dim3(kBlock), 0, st, N, nmates, occbase, mlen, wbase, words, k,
                         nb, th);
      GX_HIP(hipGetLastError());
      GX_HIP(hipMemcpyAsync(toph.data(), th, 256 * 8, hipMemcpyDeviceToHost, st));
      GX_HIP(hipStreamSynchronize(st));
      cut_ranges(0, target);
    }
  }
  uint32_t R = (uint32_t)rocc.size();
  // Several ranges that keep their results: the stores that will hold the
  // whole graph (the node arrays; the records too at P = 1) take the place
  // of the accumulators for this build and the accumulators that of the
  // per-range arrays, so a repeated build reuses one store of each size
  // (swapped back after the last range) instead of growing both to the
  // whole graph's size.
  const bool acc_swap = R > 1 && keep;
  auto swap_stores = [&] {
    swap_buf(b->klo, b->c_klo);
    if (W == 2) swap_buf(b->khi, b->c_khi);
    swap_buf(b->cnt, b->c_cnt);
    swap_buf(b->msk, b->c_msk);
    swap_buf(b->hcu, b->c_hcu);
    swap_buf(b->hcf, b->c_hcf);
    swap_buf(b->uval, b->c_uval);
    if (P == 1) swap_buf(b->out, b->outacc);
  };
  if (acc_swap) swap_stores();
  uint64_t cap = 0, rebuilds = 0, tot_nn = 0, tot_nrep = 0, tot_nsel = 0, tot_bytes = 0, acc_len = 0;
  std::vector<uint64_t> seg_src, seg_len;  // several ranges, no sink: range r's part p bytes in outacc
  uint64_t glen[7] = {};  // several ranges, no sink: bytes of the node graph so far (klo, khi, cnt, msk, heads)
  std::vector<uint64_t> prec(P + 1, 0), pbyt(P + 1, 0);  // per partition: records and bytes so far
  b->nuheads = 0;
  b->has_hoff = false;
  b->stats[9] = 0;
  // record-sort bucket bytes of a pass of n records (LDS-sized buckets)
  auto bkb_of = [&](uint64_t n) { return nb < 2 ? 1 : (nb >= 4 && n > (1ull << 28)) ? 3 : 2; };
  DevBuf *hcur = &b->selflag, *hnext = &b->hist2;  // per-tile digit counts: this range's, the next one's
  bool hist_ready = false;
  for (uint32_t ri = 0; ri < R; ri++) {
  // the ranges after this one again, from what the ranges so far measured
  // (nodes and kept bytes per occurrence) and the HBM free now; this range's
  // own bounds stay (the last pass counted its digits)
  if (!b->range_occ && ri >= 1 && ri + 1 < R) {
    uint64_t done = 0, rest = 0;
    for (uint32_t r = 0; r < ri; r++) done += rocc[r];
    for (uint32_t r = ri + 1; r < R; r++) rest += rocc[r];
    double kept_now = 0, kept_len = (double)acc_len;
    if (keep)
      for (DevBuf* d : kept_bufs) kept_now += (double)d->cap;
    for (uint64_t g : glen) kept_len += (double)g;
    const double dfrac = std::max(1e-3, (double)tot_nn / (double)std::max<uint64_t>(done, 1));
    const double kpo = keep ? kept_len / (double)std::max<uint64_t>(done, 1) : 0.0;
    const double more = std::max(0.0, 1.05 * kpo * (double)N - kept_now);  // kept bytes still to come
    cut_ranges(ri + 1, range_target(rest, pass_budget(kept_now + more), pass_bytes_per_occurrence(W, nb, dfrac)));
    R = (uint32_t)rocc.size();
  }
  const uint32_t tlo = tb[ri], thi = tb[ri + 1];
  const uint64_t Nr = rocc[ri];
  uint64_t nn = 0;
  void* table = nullptr;
  uint64_t *klo = nullptr, *khi = nullptr;
  uint32_t *cnt = nullptr, *msk = nullptr, *nslot = nullptr, *occ_slot = nullptr;
  const uint64_t* rseg = nullptr;  // sorted path: first record of each node
  const uint8_t* rbytes = nullptr;  // sorted path: record codes in node order
  uint8_t* occ_code = nullptr;
  if (sorted) {
    ph_begin(b, PH_INSERT);
    SK* ka = buf<SK>(b->ckey, Nr + 2);
    SK* kb = buf<SK>(b->ckey2, Nr + 2);
    // neighbour codes beside the keys; after the sorts the node pass keeps
    // its chunk offsets (nch + 1 words) in the free one of the pair
    const uint64_t nch = (Nr + kEmitChunk - 1) / kEmitChunk;
    const size_t vbytes = std::max<size_t>(Nr + 2, 8 * (nch + 2));
    uint8_t* va = buf<uint8_t>(b->sva, vbytes);
    uint8_t* vb = buf<uint8_t>(b->svb, vbytes);
    const unsigned ntiles = (unsigned)std::max<uint64_t>(1, (N + kSortTile - 1) / kSortTile);
    // record-sort buckets: the top bkb key bytes (LDS-sized buckets)
    const int bkb = bkb_of(Nr);
    const int dbyte = nb - bkb;
    const size_t hn = (size_t)256 * ntiles;
    uint32_t* hist = buf<uint32_t>(*hcur, hn + 1);
    uint64_t* hoff = buf<uint64_t>(b->seg, hn + 1);
    bool next_hist = false;
    if (Nr) {
      ph_begin(b, PH_KINSERT);
      if (!hist_ready) {
        hipLaunchKernelGGL(k_occ_hist<W>, dim3(ntiles), dim3(kBlock), 0, st, N, nmates, occbase, mlen, wbase, words,
                           k, nb, hist, ntiles, tlo, thi, dbyte);
        GX_HIP(hipGetLastError());
      }
      exclusive_scan<uint32_t, uint64_t>(hist, hoff, hn, scratch_for(b, scan_scratch_bytes<uint64_t>(hn)), st);
      // the same extraction pass counts the next range's digits
      next_hist = ri + 1 < R && rocc[ri + 1] > 0;
      uint32_t* nh = next_hist ? buf<uint32_t>(*hnext, hn + 1) : nullptr;
      const uint32_t nlo = next_hist ? tb[ri + 1] : 0, nhi = next_hist ? tb[ri + 2] : 0;
      const int ndbyte = next_hist ? nb - bkb_of(rocc[ri + 1]) : 0;
      ph_begin(b, PH_KSCATTER);
      hipLaunchKernelGGL(k_occ_scatter<W>, dim3(ntiles), dim3(kBlock), 0, st, N, nmates, occbase, mlen, wbase, words,
                         k, nb, hoff, ntiles, ka, va, tlo, thi, dbyte, nlo, nhi, ndbyte, nh);
      GX_HIP(hipGetLastError());
      ph_end(b, PH_KSCATTER);
      ph_end(b, PH_KINSERT);
    }
    hist_ready = next_hist;
    std::swap(hcur, hnext);
    // the more significant bucket bytes
    SK *kx = ka, *ky = kb;
    uint8_t *vx = va, *vy = vb;
    if (bkb >= 2 && Nr > 1) {
      bool f = radix_sort<SK, uint8_t>(ka, va, kb, vb, Nr, nb - bkb + 1, nb,
                                        scratch_for(b, radix_scratch_bytes<SK>(Nr)), st);
      if (f) { std::swap(kx, ky); std::swap(vx, vy); }
    }
    // buckets: LDS sort, oversized ones by a global radix sort; the blocks
    // cover the buckets of this key range only
    const uint32_t NBK = 1u << (8 * bkb);
    const uint32_t bk0 = tlo << (8 * (bkb - 1)), bk1 = thi << (8 * (bkb - 1));
    uint64_t* bstart = buf<uint64_t>(b->hstart, NBK + 2);
    uint32_t* big = buf<uint32_t>(b->hcu, NBK + 1);
    uint64_t* bsz = buf<uint64_t>(b->htot, NBK + 1);
    uint64_t* boff = buf<uint64_t>(b->upos, NBK + 2);
    GX_HIP(hipMemsetAsync(big, 0, (NBK + 1) * 4, st));
    hipLaunchKernelGGL(k_bucket_starts<SK>, dim3(grid_for(Nr + 1, kBlock, G)), dim3(kBlock), 0, st, Nr, kx, nb, bkb,
                       bk0, bk1, bstart);
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemsetAsync(errw + 2, 0xff, 8, st));
    // sorted records: keys in ky, neighbour codes (the replay's code array) in rcodes
    uint8_t* rcodes = buf<uint8_t>(b->rcode, Nr + 32);
    uint32_t* blists = buf<uint32_t>(b->blist, 3 * (size_t)NBK + 3);
    uint32_t* nlist = blists + 3 * (size_t)NBK;
    GX_HIP(hipMemsetAsync(nlist, 0, 12, st));
    const unsigned kMaxSortBlocks = 1u << 20;
    hipLaunchKernelGGL((k_bucket_sort<W, kSmallBucket, 64>), dim3(std::min(bk1 - bk0, kMaxSortBlocks)), dim3(64), 0,
                       st, kx, vx, ky, rcodes, bstart, Nr, nb, bkb, bk0, bk1 - bk0, nullptr, big, blists, nlist,
                       NBK, errw + 2);
    GX_HIP(hipGetLastError());
    uint32_t nl[3];
    GX_HIP(hipMemcpyAsync(nl, nlist, 12, hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    if (nl[0] > bk1 - bk0 || nl[1] > bk1 - bk0 || nl[2] > bk1 - bk0)
      throw GxError(GX_EDEVICE, "internal error: bucket list overflow");
    if (nl[0]) {
      hipLaunchKernelGGL((k_bucket_sort<W, kBucketCap / 8, 256>), dim3(std::min(nl[0], 65536u)), dim3(256), 0, st, kx,
                         vx, ky, rcodes, bstart, Nr, nb, bkb, bk0, nl[0], blists, big, nullptr, nullptr, NBK, errw + 2);
      GX_HIP(hipGetLastError());
    }
    if (nl[1]) {
      hipLaunchKernelGGL((k_bucket_sort<W, kBucketCap / 2, 512>), dim3(std::min(nl[1], 65536u)), dim3(512), 0, st, kx,
                         vx, ky, rcodes, bstart, Nr, nb, bkb, bk0, nl[1], blists + NBK, big, nullptr, nullptr, NBK,
                         errw + 2);
      GX_HIP(hipGetLastError());
    }
    if (nl[2]) {
      hipLaunchKernelGGL((k_bucket_sort<W, kBucketCap, 1024>), dim3(std::min(nl[2], 65536u)), dim3(1024), 0, st, kx,
                         vx, ky, rcodes, bstart, Nr, nb, bkb, bk0, nl[2], blists + 2 * (size_t)NBK, big, nullptr,
                         nullptr, NBK, errw + 2);
      GX_HIP(hipGetLastError());
    }
    if (read1(st, errw + 2) != ~0ull) throw GxError(GX_EDEVICE, "internal error: unordered bucket bounds or lists");
    hipLaunchKernelGGL(k_big_sizes, dim3(grid_for(NBK, kBlock, G)), dim3(kBlock), 0, st, NBK, bstart, big, bsz);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint64_t, uint64_t>(bsz, boff, NBK, scratch_for(b, scan_scratch_bytes<uint64_t>(NBK)), st);
    const uint64_t nbig = read1(st, boff + NBK);
    if (nbig) {
      SK* tk = buf<SK>(b->hkey, nbig + 1);
      SK* tk2 = buf<SK>(b->hkey2, nbig + 1);
      uint8_t* tv = buf<uint8_t>(b->hval, nbig + 1);
      uint8_t* tv2 = buf<uint8_t>(b->hval2, nbig + 1);
      hipLaunchKernelGGL(k_big_move<SK>, dim3(std::min(bk1 - bk0, 65536u)), dim3(kBlock), 0, st, bstart, big, boff, kx,
                         vx, ky, rcodes, tk, tv, 0, bk0, bk1);
      GX_HIP(hipGetLastError());
      // all key bytes: several oversized buckets share the array, and the
      // stable sort keeps each bucket's records in occurrence order
      bool f = radix_sort<SK, uint8_t>(tk, tv, tk2, tv2, nbig, 0, nb, scratch_for(b, radix_scratch_bytes<SK>(nbig)),
                                        st);
      if (f) { std::swap(tk, tk2); std::swap(tv, tv2); }
      hipLaunchKernelGGL(k_big_move<SK>, dim3(std::min(bk1 - bk0, 65536u)), dim3(kBlock), 0, st, bstart, big, boff, kx,
                         vx, ky, rcodes, tk, tv, 1, bk0, bk1);
      GX_HIP(hipGetLastError());
    }
    b->stats[9] += nbig;
    ph_end(b, PH_INSERT);

    // 5'. nodes = runs of equal keys
    ph_begin(b, PH_NODES);
    // run starts per chunk of kEmitChunk records, and each chunk's first node
    uint32_t* ccnt = reinterpret_cast<uint32_t*>(kx);  // the other buffers are free now
    uint64_t* coff = reinterpret_cast<uint64_t*>(vx);
    if (Nr) {
      hipLaunchKernelGGL(k_run_count<SK>, dim3(grid_for(nch, 1, 1u << 16)), dim3(kBlock), 0, st, Nr, ky, ccnt);
      GX_HIP(hipGetLastError());
    }
    exclusive_scan<uint32_t, uint64_t>(ccnt, coff, nch, scratch_for(b, scan_scratch_bytes<uint64_t>(nch)), st);
    nn = read1(st, coff + nch);
    if (nn >= 0xffffffffull) throw GxError(GX_ENOTSUP, "more than 2^32-1 distinct k-mers");
    klo = buf<uint64_t>(b->klo, nn + 1);
    khi = W == 2 ? buf<uint64_t>(b->khi, nn + 1) : nullptr;
    cnt = buf<uint32_t>(b->cnt, nn + 1);
    msk = buf<uint32_t>(b->msk, nn + 1);
    uint64_t* nst = buf<uint64_t>(b->nslot, nn + 2);
    GX_HIP(hipMemsetAsync(msk, 0, (nn + 1) * 4, st));
    if (Nr) {
      hipLaunchKernelGGL(k_node_emit<W>, dim3(grid_for(nch, 1, 1u << 16)), dim3(kBlock), 0, st, Nr, ky, rcodes, coff,
                         nb, klo, khi, nst, msk);
      GX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, st, nst + nn, Nr);
    GX_HIP(hipGetLastError());
    if (nn) {
      hipLaunchKernelGGL(k_node_cnt, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, nst, cnt);
      GX_HIP(hipGetLastError());
    }
    rseg = nst;
    rbytes = rcodes;
    b->reclaim.dead.insert(b->reclaim.dead.end(), {&b->ckey, &b->ckey2, &b->sva, &b->svb});  // records are dead
    ph_end(b, PH_NODES);
  } else {
    // 4. table insert (grow-and-rebuild on overflow).  Capacity: the next power
    //    of two >= the expected distinct count (table_hint) or the occurrence
    //    count, so the load stays below 1 and is typically well below 1/2.
    ph_begin(b, PH_INSERT);
    cap = learned_cap;
    if (!cap) {
      uint64_t expect = b->cfg.table_hint ? b->cfg.table_hint * 2 : N;
      cap = 1024;
      while (cap < expect) cap <<= 1;
    }
    occ_slot = buf<uint32_t>(b->occ_slot, N + 1);
    occ_code = buf<uint8_t>(b->occ_code, N + 1);
    for (int attempt = 0;; attempt++) {
      if (cap > (1ull << 32)) throw GxError(GX_ENOTSUP, "hash table beyond 2^32 slots");
      table = b->table.ensure(cap * slotsz);
      hipLaunchKernelGGL(k_table_clear<W>, dim3(grid_for(cap, kBlock, 65536)), dim3(kBlock), 0, st, table, cap);
      GX_HIP(hipGetLastError());
      GX_HIP(hipMemsetAsync(errw, 0xff, 8, st));
      // probe bound: far above the longest probe of a table at load <= 0.7;
      // a nearly full (under-sized) table fails fast and is rebuilt larger
      uint64_t limit = std::min<uint64_t>(cap, 4096);
      if (nmates) {
        unsigned g = grid_for(nmates, kBlock / 64, 1u << 16);
        ph_begin(b, PH_KINSERT);
        hipLaunchKernelGGL(k_insert<W>, dim3(g), dim3(kBlock), 0, st, nmates, mlen, wbase, words, occbase, nocc, k,
                           table, cap, limit, occ_slot, occ_code, errw);
        GX_HIP(hipGetLastError());
        ph_end(b, PH_KINSERT);
      }
      unsigned long long e = read1(st, errw);
      if (e == ~0ull) break;
      if ((e & 0xff) != ERR_TABLE || attempt > 8) check_err(b);
      // fallback: rebuild in a table twice as large, and at least as large as
      // the sizing from the occurrence count (the method without a prediction)
      uint64_t plain = 1024;
      while (plain < N) plain <<= 1;
      cap = std::max(cap << 1, rebuilds == 0 ? plain : 0);
      rebuilds++;
    }
    ph_end(b, PH_INSERT);

    // 5. KmerPointable order
    ph_begin(b, PH_NODES);
    unsigned cblocks = grid_for(cap, kCTile, 0xffffffffu);
    uint64_t* cb = buf<uint64_t>(b->blk, (size_t)cblocks + 1);
    uint64_t* cbb = buf<uint64_t>(b->blk2, (size_t)cblocks + 1);
    hipLaunchKernelGGL(k_compact_count<W>, dim3(cblocks), dim3(kBlock), 0, st, table, cap, cb);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint64_t, uint64_t>(cb, cbb, cblocks, scratch_for(b, scan_scratch_bytes<uint64_t>(cblocks)), st);
    nn = read1(st, cbb + cblocks);
    if (nn >= 0xffffffffull) throw GxError(GX_ENOTSUP, "more than 2^32-1 distinct k-mers");
    uint32_t* claimed = buf<uint32_t>(b->claimed, nn + 1);
    hipLaunchKernelGGL(k_compact_write<W>, dim3(cblocks), dim3(kBlock), 0, st, table, cap, cbb, claimed);
    GX_HIP(hipGetLastError());
    SK* ck = buf<SK>(b->ckey, nn + 1);
    SK* ck2 = buf<SK>(b->ckey2, nn + 1);
    uint32_t* cv = buf<uint32_t>(b->cval, nn + 1);
    uint32_t* cv2 = buf<uint32_t>(b->cval2, nn + 1);
    if (nn) {
      hipLaunchKernelGGL((k_compact_list<W, SK>), dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, table, claimed,
                         nn, nb, ck, cv);
      GX_HIP(hipGetLastError());
    }
    bool fl = radix_sort<SK>(ck, cv, ck2, cv2, nn, 0, nb, scratch_for(b, radix_scratch_bytes<SK>(nn)), st);
    uint32_t* perm = fl ? cv2 : cv;
    klo = buf<uint64_t>(b->klo, nn + 1);
    khi = W == 2 ? buf<uint64_t>(b->khi, nn + 1) : nullptr;
    cnt = buf<uint32_t>(b->cnt, nn + 1);
    msk = buf<uint32_t>(b->msk, nn + 1);
    nslot = buf<uint32_t>(b->nslot, nn + 1);
    if (nn) {
      hipLaunchKernelGGL(k_gather_nodes<W>, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, perm, claimed, table,
                         klo, khi, cnt, msk, nslot);
      GX_HIP(hipGetLastError());
      GX_HIP(hipMemsetAsync(errw + 2, 0xff, 8, st));
      hipLaunchKernelGGL(k_dup_check<W>, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, klo, khi, errw + 2);
      GX_HIP(hipGetLastError());
      unsigned long long dup = read1(st, errw + 2);
      if (dup != ~0ull) throw GxError(GX_EDEVICE, "internal error: k-mer stored in two table slots (node " +
                                                      std::to_string(dup) + ")");
    }
    b->reclaim.dead.insert(b->reclaim.dead.end(), {&b->ckey, &b->ckey2, &b->cval, &b->cval2, &b->claimed});
    ph_end(b, PH_NODES);

  }

  // 5b. key -> node index (heads, unitig links)
  ph_begin(b, PH_NIX);
  const NodeIndex nix = build_node_index<W>(b, st, nn, klo, khi, msk, tlo, thi);
  if (sorted) cap = 2 * nn;
  ph_end(b, PH_NIX);
  bool side_open = false;  // k_replay_long runs on the side stream, not joined yet

  // 6. replay
  ph_begin(b, PH_REPLAY);
  const int jdk7 = b->cfg.edge_order == GX_EDGE_ORDER_JDK7;
  uint32_t* needs = buf<uint32_t>(b->needs, nn + 1);
  uint64_t* selcnt = buf<uint64_t>(b->selcnt, nn + 1);
  uint64_t* seg = buf<uint64_t>(b->seg, nn + 1);
  if (nn) {
    hipLaunchKernelGGL(k_node_info<W>, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, msk, cnt, nslot,
                       sorted ? nullptr : table, jdk7, needs, selcnt, errw + 3);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(selcnt, seg, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
  uint64_t nsel = read1(st, seg + nn);
  uint64_t* lists = buf<uint64_t>(b->lists, nn + 1);
  uint64_t nrep = 0;
  if (nsel) {
    const uint32_t* codes = nullptr;
    if (!sorted) {
    uint32_t* flag = buf<uint32_t>(b->selflag, N + 1);
    uint64_t* pos = buf<uint64_t>(b->selpos, N + 1);
    hipLaunchKernelGGL(k_sel_flags<W>, dim3(grid_for(N, kBlock, G)), dim3(kBlock), 0, st, N, occ_slot, table, flag);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint32_t, uint64_t>(flag, pos, N, scratch_for(b, scan_scratch_bytes<uint64_t>(N)), st);
    uint32_t* sk = buf<uint32_t>(b->selk, nsel + 1);
    uint32_t* sv = buf<uint32_t>(b->selv, nsel + 1);
    uint32_t* sk2 = buf<uint32_t>(b->selk2, nsel + 1);
    uint32_t* sv2 = buf<uint32_t>(b->selv2, nsel + 1);
    hipLaunchKernelGGL(k_sel_write<W>, dim3(grid_for(N, kBlock, G)), dim3(kBlock), 0, st, N, occ_slot, occ_code, table,
                       flag, pos, sk, sv);
    GX_HIP(hipGetLastError());
    int nbytes_id = 1;
    while (nbytes_id < 4 && (nn >> (8 * nbytes_id)) > 0) nbytes_id++;
    bool f2 = radix_sort<uint32_t>(sk, sv, sk2, sv2, nsel, 0, nbytes_id,
                                   scratch_for(b, radix_scratch_bytes<uint32_t>(nsel)), st);
    codes = f2 ? sv2 : sv;
    uint8_t* cb = buf<uint8_t>(b->rcode, nsel + 32);
    hipLaunchKernelGGL(k_narrow_codes, dim3(grid_for(nsel, kBlock, G)), dim3(kBlock), 0, st, nsel, codes, cb);
    GX_HIP(hipGetLastError());
    rbytes = cb;
    rseg = seg;
    // occurrence records, selection and table are dead now
    b->reclaim.dead.insert(b->reclaim.dead.end(), {&b->occ_slot, &b->occ_code, &b->selflag, &b->selpos, &b->selk,
                                                   &b->selv, &b->selk2, &b->selv2, &b->table});
    table = nullptr;
    }
    // replay list ordered by descending record count
    uint64_t* rpos = buf<uint64_t>(b->rpos, nn + 1);
    exclusive_scan<uint32_t, uint64_t>(needs, rpos, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
    nrep = read1(st, rpos + nn);
    uint32_t* rk = buf<uint32_t>(b->rkey, nrep + 1);
    uint32_t* rv = buf<uint32_t>(b->rval, nrep + 1);
    uint32_t* rk2 = buf<uint32_t>(b->rkey2, nrep + 1);
    uint32_t* rv2 = buf<uint32_t>(b->rval2, nrep + 1);
    hipLaunchKernelGGL(k_replay_list, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, needs, rpos, cnt, rk, rv);
    GX_HIP(hipGetLastError());
    bool f4 = radix_sort<uint32_t>(rk, rv, rk2, rv2, nrep, 0, 2, scratch_for(b, radix_scratch_bytes<uint32_t>(nrep)),
                                   st);
    uint32_t* rlist = f4 ? rv2 : rv;
    const uint32_t* rkeys = f4 ? rk2 : rk;
    // the longest runs (a prefix of the list) go to the four-lane kernel
    GX_HIP(hipMemsetAsync(errw + 2, 0, 8, st));
    hipLaunchKernelGGL(k_count_long, dim3(grid_for(nrep, kBlock, G)), dim3(kBlock), 0, st, nrep, rkeys,
                       0xffffu - kLongReplay, errw + 2);
    GX_HIP(hipGetLastError());
    const uint64_t nlong = read1(st, errw + 2);
    ph_begin(b, PH_KREPLAY);
    if (nlong) {
      // a few hundred waves, each as long as its longest run: on the side
      // stream, overlapping the short runs, the heads and the record sizes;
      // joined before the records are written
      side_init(b);
      stream_after(b->st2, st, b->ev_fork);
      side_open = true;
      hipLaunchKernelGGL(k_replay_long<W>, dim3(grid_for(4 * nlong, kBlock, 1u << 16)), dim3(kBlock), 0, b->st2,
                         nlong, rlist, rseg, rbytes, klo, khi, msk, k, nb, lists);
      GX_HIP(hipGetLastError());
    }
    if (nrep > nlong) {
      hipLaunchKernelGGL(k_replay<W>, dim3(grid_for(nrep - nlong, kBlock, 1u << 16)), dim3(kBlock), 0, st,
                         nrep - nlong, rlist + nlong, rseg, rbytes, klo, khi, msk, k, nb, lists);
      GX_HIP(hipGetLastError());
    }
    ph_end(b, PH_KREPLAY);
  }
  ph_end(b, PH_REPLAY);
  // replay inputs and order are dead (lists, needs and the node arrays
  // stay); a free to make room drains the side stream first
  b->reclaim.dead.insert(b->reclaim.dead.end(), {&b->nslot, &b->rcode, &b->selcnt, &b->seg, &b->rpos, &b->rkey,
                                                 &b->rval, &b->rkey2, &b->rval2});

  // 7. heads
  ph_begin(b, PH_HEADS);
  uint64_t* hk = buf<uint64_t>(b->hkey, nheads + 1);
  uint32_t* hv = buf<uint32_t>(b->hval, nheads + 1);
  uint64_t* hk2 = buf<uint64_t>(b->hkey2, nheads + 1);
  uint32_t* hv2 = buf<uint32_t>(b->hval2, nheads + 1);
  if (nmates) {
    hipLaunchKernelGGL(k_heads<W>, dim3(grid_for(nmates, kBlock, G)), dim3(kBlock), 0, st, nmates, valid, nocc, hidx,
                       nix, rid, wbase, words, k, hk, hv, tlo, thi);
    GX_HIP(hipGetLastError());
  }
  uint32_t* uflag = buf<uint32_t>(b->uflag, nheads + 1);
  uint64_t* upos = buf<uint64_t>(b->upos, nheads + 1);
  uint64_t nh = nheads;  // heads of this key range
  if (R > 1 && nheads) {  // keep this range's heads (in mate order) before sorting
    hipLaunchKernelGGL(k_head_in_range, dim3(grid_for(nheads, kBlock, G)), dim3(kBlock), 0, st, nheads, hk, uflag);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint32_t, uint64_t>(uflag, upos, nheads, scratch_for(b, scan_scratch_bytes<uint64_t>(nheads)), st);
    nh = read1(st, upos + nheads);
    hipLaunchKernelGGL(k_head_keep, dim3(grid_for(nheads, kBlock, G)), dim3(kBlock), 0, st, nheads, hk, hv, uflag,
                       upos, hk2, hv2);
    GX_HIP(hipGetLastError());
    std::swap(hk, hk2);
    std::swap(hv, hv2);
  }
  int hbytes = 1;
  while (hbytes < 8 && (((nn << 31) | ((1ull << 31) - 1)) >> (8 * hbytes)) > 0) hbytes++;
  bool f3 = radix_sort<uint64_t>(hk, hv, hk2, hv2, nh, 0, hbytes, scratch_for(b, radix_scratch_bytes<uint64_t>(nh)),
                                 st);
  if (f3) { std::swap(hk, hk2); std::swap(hv, hv2); }
  uint32_t* uval = buf<uint32_t>(b->uval, nheads + 1);
  uint32_t* hcu = buf<uint32_t>(b->hcu, nn + 1);
  uint32_t* hcf = buf<uint32_t>(b->hcf, nn + 1);
  uint64_t* htot = buf<uint64_t>(b->htot, nn + 1);
  uint64_t* hstart = buf<uint64_t>(b->hstart, nn + 1);
  GX_HIP(hipMemsetAsync(hcu, 0, (nn + 1) * 4, st));
  GX_HIP(hipMemsetAsync(hcf, 0, (nn + 1) * 4, st));
  if (nh) {
    hipLaunchKernelGGL(k_head_unique, dim3(grid_for(nh, kBlock, G)), dim3(kBlock), 0, st, nh, hk, uflag);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint32_t, uint64_t>(uflag, upos, nh, scratch_for(b, scan_scratch_bytes<uint64_t>(nh)), st);
  const uint64_t nuh_r = read1(st, upos + nh);  // unique heads of this range
  b->nuheads += nuh_r;
  if (nh) {
    hipLaunchKernelGGL(k_head_compact, dim3(grid_for(nh, kBlock, G)), dim3(kBlock), 0, st, nh, hk, hv, uflag, upos,
                       uval, hcu, hcf);
    GX_HIP(hipGetLastError());
  }
  if (nn) {
    hipLaunchKernelGGL(k_head_total, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, hcu, hcf, htot);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(htot, hstart, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
  ph_end(b, PH_HEADS);
  // head sort arrays are dead; with several key ranges so is the range's
  // node index (rebuilt over the whole graph after the last range)
  b->reclaim.dead.insert(b->reclaim.dead.end(), {&b->hkey, &b->hval, &b->hkey2, &b->hval2, &b->uflag, &b->upos});
  if (R > 1) b->reclaim.dead.push_back(&b->nix_slots);

  // 8. serialisation
  ph_begin(b, PH_SER);
  SerCtx c;
  c.k = k; c.nb = nb; c.jdk7 = jdk7; c.nn = nn;
  c.klo = klo; c.khi = khi; c.cnt = cnt; c.msk = msk; c.needs = needs; c.lists = lists;
  c.hstart = hstart; c.hcu = hcu; c.hcf = hcf; c.uval = uval;
  c.mlen = mlen; c.wbase = wbase; c.words = words; c.rid = rid;
  c.bad = errw + 1;
  c.perm = nullptr;
  GX_HIP(hipMemsetAsync(errw + 1, 0xff, 8, st));
  // P > 1: records are grouped by output partition, each group in key order
  uint64_t* pstart = nullptr;
  if (P > 1) {
    uint32_t* pid = buf<uint32_t>(b->ppid, nn + 1);
    uint32_t* pidx = buf<uint32_t>(b->pidx, nn + 1);
    uint32_t* pid2 = buf<uint32_t>(b->ppid2, nn + 1);
    uint32_t* pidx2 = buf<uint32_t>(b->pidx2, nn + 1);
    pstart = buf<uint64_t>(b->pbound, 2 * ((uint64_t)P + 1));
    if (nn) {
      hipLaunchKernelGGL(k_part_ids<W>, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, klo, khi, nb, P, pid,
                         pidx);
      GX_HIP(hipGetLastError());
    }
    int pbytes = 1;
    while (pbytes < 4 && ((P - 1) >> (8 * pbytes)) > 0) pbytes++;
    bool fp = radix_sort<uint32_t>(pid, pidx, pid2, pidx2, nn, 0, pbytes,
                                   scratch_for(b, radix_scratch_bytes<uint32_t>(nn)), st);
    if (fp) { std::swap(pid, pid2); std::swap(pidx, pidx2); }
    hipLaunchKernelGGL(k_part_starts, dim3(grid_for(nn + 1, kBlock, G)), dim3(kBlock), 0, st, nn, pid, P, pstart);
    GX_HIP(hipGetLastError());
    c.perm = pidx;
  }
  uint64_t* rsize = buf<uint64_t>(b->rsize, nn + 1);
  uint64_t* roff = buf<uint64_t>(b->roff, nn + 1);
  if (nn) {
    hipLaunchKernelGGL(k_ser_size, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, c, rsize);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(rsize, roff, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
  uint64_t total = read1(st, roff + nn);
  b->reclaim.dead.push_back(&b->rsize);
  b->part_rec.assign(P + 1, 0);
  b->part_off.assign(P + 1, 0);
  if (P > 1) {
    uint64_t* pbyte = pstart + P + 1;
    hipLaunchKernelGGL(k_part_bytes, dim3(grid_for((uint64_t)P + 1, kBlock, G)), dim3(kBlock), 0, st, P, pstart, roff,
                       pbyte);
    GX_HIP(hipGetLastError());
    GX_HIP(hipMemcpyAsync(b->part_rec.data(), pstart, (P + 1) * 8, hipMemcpyDeviceToHost, st));
    GX_HIP(hipMemcpyAsync(b->part_off.data(), pbyte, (P + 1) * 8, hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
  } else {
    b->part_rec[1] = nn;
    b->part_off[1] = total;
  }
  if (side_open) {  // the records need every replayed list
    stream_after(st, b->st2, b->ev_join);
    side_open = false;
  }
  uint8_t* out = buf<uint8_t>(b->out, total + 1);
  if (nn) {
    hipLaunchKernelGGL(k_ser_write<W>, dim3(grid_for(nn, kBlock, 0xffffffffu)), dim3(kBlock), 0, st, c, roff, out);
    GX_HIP(hipGetLastError());
  }
  ph_end(b, PH_SER);
  GX_HIP(hipStreamSynchronize(st));
  ph_collect(b, true);
  unsigned long long badnode = read1(st, errw + 1);
  if (badnode != ~0ull)
    throw GxError(GX_EDEVICE, "internal error: replayed edge list inconsistent at node " + std::to_string(badnode));
  tot_nn += nn;
  tot_nrep += nrep;
  tot_nsel += nsel;
  tot_bytes += total;
  b->nnodes = nn;
  if (b->sink) {  // the range's records, partition by partition, to the caller
    for (uint32_t p = 0; p < P; p++) {
      const uint64_t o = b->part_off[p], nbyt = b->part_off[p + 1] - o, nrec = b->part_rec[p + 1] - b->part_rec[p];
      if (!nbyt) continue;
      int rc = b->sink(b->sink_user, (int)p, out + o, nbyt, nrec);
      if (rc) throw GxError(rc < 0 ? rc : GX_EINVAL, "sink returned " + std::to_string(rc));
    }
  } else if (R > 1) {  // kept in HBM; partitions are assembled after the last range
    // the first range sizes each store for the whole build from its own
    // bytes per occurrence (no re-allocation with both copies live later)
    auto whole = [&](size_t len) -> size_t {
      return ri == 0 && Nr * 256 >= N ? (size_t)(1.15 * (double)len * (double)N / (double)Nr) : 0;
    };
    append_to(b, b->outacc, acc_len, out, total, hipMemcpyDeviceToDevice, whole(total));
    for (uint32_t p = 0; p < P; p++) {
      seg_src.push_back(acc_len - total + b->part_off[p]);
      seg_len.push_back(b->part_off[p + 1] - b->part_off[p]);
    }
    // and the range's node graph: ranges follow key order, so the nodes of
    // range after range are the whole graph's nodes in key order; heads are
    // stored per node and carry no node id
    const hipMemcpyKind dd = hipMemcpyDeviceToDevice;
    append_to(b, b->c_klo, glen[0], klo, nn * 8, dd, whole(nn * 8));
    if (W == 2) append_to(b, b->c_khi, glen[1], khi, nn * 8, dd, whole(nn * 8));
    append_to(b, b->c_cnt, glen[2], cnt, nn * 4, dd, whole(nn * 4));
    append_to(b, b->c_msk, glen[3], msk, nn * 4, dd, whole(nn * 4));
    append_to(b, b->c_hcu, glen[4], hcu, nn * 4, dd, whole(nn * 4));
    append_to(b, b->c_hcf, glen[5], hcf, nn * 4, dd, whole(nn * 4));
    append_to(b, b->c_uval, glen[6], uval, nuh_r * 4, dd, whole(nuh_r * 4));
  }
  for (uint32_t p = 0; p < P; p++) {
    prec[p + 1] += b->part_rec[p + 1] - b->part_rec[p];
    pbyt[p + 1] += b->part_off[p + 1] - b->part_off[p];
  }
  if (R > 1)  // the next range takes these buffers back as it needs them
    b->reclaim.dead.insert(b->reclaim.dead.end(),
                           {&b->klo, &b->khi, &b->cnt, &b->msk, &b->nslot, &b->nix_slots, &b->needs, &b->selcnt,
                            &b->seg, &b->lists, &b->rpos, &b->rkey, &b->rval, &b->rkey2, &b->rval2, &b->rcode,
                            &b->hcu, &b->hcf, &b->htot, &b->hstart, &b->uval, &b->rsize, &b->roff, &b->out, &b->ppid,
                            &b->pidx, &b->ppid2, &b->pidx2});
  }  // key ranges

  for (uint32_t p = 0; p < P; p++) {
    prec[p + 1] += prec[p];
    pbyt[p + 1] += pbyt[p];
  }
  if (!b->sink && R > 1) {
    // part p = its bytes of range 0, then of range 1, ...
    if (P == 1) {
      swap_buf(b->out, b->outacc);
      b->out.ensure(tot_bytes + 1);  // live again: off the reclaim list (no allocation)
    } else {
      uint8_t* out = buf<uint8_t>(b->out, tot_bytes + 1);
      uint64_t dst = 0;
      for (uint32_t p = 0; p < P; p++)
        for (uint32_t r = 0; r < R; r++) {
          const size_t i = (size_t)r * P + p;
          if (seg_len[i])
            GX_HIP(hipMemcpyAsync(out + dst, b->outacc.as<uint8_t>() + seg_src[i], seg_len[i],
                                  hipMemcpyDeviceToDevice, st));
          dst += seg_len[i];
        }
      GX_HIP(hipStreamSynchronize(st));
    }
    b->reclaim.dead.push_back(&b->outacc);
    // the whole node graph becomes the handle's, with its heads' starts and
    // a node index over every node (the graph steps look nodes up)
    DevBuf* held[7] = {&b->klo, &b->khi, &b->cnt, &b->msk, &b->hcu, &b->hcf, &b->uval};
    DevBuf* acc[7] = {&b->c_klo, &b->c_khi, &b->c_cnt, &b->c_msk, &b->c_hcu, &b->c_hcf, &b->c_uval};
    for (int i = 0; i < 7; i++) {
      if (i == 1 && W != 2) continue;
      swap_buf(*held[i], *acc[i]);
      held[i]->ensure(glen[i] + 1);  // live again: off the reclaim list (no allocation)
      b->reclaim.dead.push_back(acc[i]);
    }
    const uint64_t nn = tot_nn;
    if (nn < 0xffffffffull) {  // the node index and the graph steps number nodes in 32 bits
      uint64_t* htot = buf<uint64_t>(b->htot, nn + 1);
      uint64_t* hstart = buf<uint64_t>(b->hstart, nn + 1);
      if (nn) {
        hipLaunchKernelGGL(k_head_total, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn,
                           b->hcu.as<uint32_t>(), b->hcf.as<uint32_t>(), htot);
        GX_HIP(hipGetLastError());
      }
      exclusive_scan<uint64_t, uint64_t>(htot, hstart, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
      ph_begin(b, PH_NIX);
      build_node_index<W>(b, b->st, nn, b->klo.as<uint64_t>(), W == 2 ? b->khi.as<uint64_t>() : nullptr,
                          b->msk.as<uint32_t>(), 0u, 256u);
      ph_end(b, PH_NIX);
      GX_HIP(hipStreamSynchronize(st));
      ph_collect(b, true);
    }
    b->nnodes = nn;
  }
  b->graph_held = R == 1 || (!b->sink && tot_nn < 0xffffffffull);
  b->part_rec = prec;
  b->part_off = pbyt;
  b->streamed = b->sink != nullptr;
  b->nranges = R;
  b->out_bytes = tot_bytes;
  b->stats[0] = N;
  b->stats[1] = tot_nn;
  b->stats[2] = nheads;
  b->stats[3] = cap;
  b->stats[4] = tot_nrep;
  b->stats[5] = tot_nsel;
  b->stats[6] = read1(st, errw + 3);
  b->stats[7] = (uint64_t)b->pred[0];
  b->stats[8] = rebuilds;
  b->stats[11] = R;
}

template <int W>
static void run_unitigs(gx_build* b) {
  hipStream_t st = b->st;
  const int k = b->cfg.k, nb = b->nb;
  const uint64_t nn = b->nnodes;
  const unsigned G = b->grid;
  const NodeIndex nix{b->nix_dir.as<uint32_t>(), b->nix_slots.p, nb, 1u << (8 * b->nix_bkb), b->nix_bkb};
  const uint64_t* klo = b->klo.as<uint64_t>();
  const uint64_t* khi = W == 2 ? b->khi.as<uint64_t>() : nullptr;
  const uint32_t* cnt = b->cnt.as<uint32_t>();
  const uint32_t* msk = b->msk.as<uint32_t>();
  unsigned long long* errw = buf<unsigned long long>(b->errw, 4);
  GX_HIP(hipMemsetAsync(errw, 0xff, 16, st));
  GX_HIP(hipMemsetAsync(errw + 2, 0, 8, st));  // successor reads of the pointer-jumping rounds
  if (nn >= kMaxUnitigNodes)
    throw GxError(GX_ENOTSUP, "the unitig pass holds fewer than 2^30 distinct k-mers; this graph has " +
                                  std::to_string(nn));
  uint64_t ne = 2 * nn;
  ph_begin(b, PH_UNITIG);
  uint32_t* lnk = buf<uint32_t>(b->u_lnk, ne + 1);
  uint8_t* tand = buf<uint8_t>(b->u_tand, nn + 1);
  uint8_t* cut = buf<uint8_t>(b->u_cut, ne + 1);
  GX_HIP(hipMemsetAsync(cut, 0, ne + 1, st));
  // ranking by walks (odd k): k_links writes the walk entries too
  const bool walk = nn && (k & 1) && !b->no_walk;
  uint4* E = walk ? buf<uint4>(b->u_ptr[0], ne + 1) : nullptr;  // the pointer-jumping buffers, free until a fallback
  uint32_t* endp = walk ? buf<uint32_t>(b->u_walk, nn + 1) : nullptr;
  // odd k: each link looked up once, by the node of the smaller key
  const bool half = (k & 1) != 0;
  if (half) GX_HIP(hipMemsetAsync(lnk, 0xff, (ne + 1) * 4, st));
  if (nn) {
    hipLaunchKernelGGL(k_links<W>, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, klo, khi, msk, nix, k, half,
                       lnk, tand, errw, half ? nullptr : E, endp);
    GX_HIP(hipGetLastError());
    if (half && E) {
      hipLaunchKernelGGL(k_link_ends, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, lnk, E, endp);
      GX_HIP(hipGetLastError());
    }
  }
  uint32_t* urep = buf<uint32_t>(b->u_rep, nn + 1);
  uint32_t* upos = buf<uint32_t>(b->u_pos, nn + 1);
  uint32_t* isrep = buf<uint32_t>(b->u_isrep, nn + 1);
  uint64_t* ulen_of_rep = buf<uint64_t>(b->u_lenrep, nn + 1);
  // ranking by walks (odd k), pointer jumping when a walk cannot rank every node
  bool walked = false;
  b->u_walkers = 0;
  if (walk) {
    ph_begin(b, PH_KWALK);
    uint64_t* at = buf<uint64_t>(b->u_uid, nn + 1);
    uint32_t* ws = buf<uint32_t>(b->u_any, 4);  // [0] capped, [1] unranked nodes
    GX_HIP(hipMemsetAsync(ws, 0, 16, st));
    const unsigned gn = grid_for(nn, kBlock, G * 4);
    exclusive_scan<uint32_t, uint64_t>(endp, at, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
    const uint64_t nw = read1(st, at + nn);
    uint32_t* walkers = buf<uint32_t>(b->u_walk2, nw + 1);
    uint64_t* sum = buf<uint64_t>(b->u_ptr[1], 3 * (nw + 1));  // path sums, walker claims, meeting records
    uint64_t* claim = sum + (nw + 1);
    uint64_t* met = claim + (nw + 1);
    GX_HIP(hipMemsetAsync(sum, 0xff, (nw + 1) * 8, st));  // a sum no join writes sends the pass to pointer jumping
    if (nw) {
      hipLaunchKernelGGL(k_walk_list, dim3(gn), dim3(kBlock), 0, st, nn, endp, at, E, walkers);
      GX_HIP(hipGetLastError());
      // one resident wave set, each lane walking ~a dozen paths in turn: a
      // wave then waits for the longest of its lanes' sums, not of single paths
      hipLaunchKernelGGL(k_walk, dim3(grid_for(nw, kBlock, G)), dim3(kBlock), 0, st, (uint32_t)nw, walkers, E, claim,
                         met, ws);
      GX_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_walk_join, dim3(grid_for(nw, kBlock, G)), dim3(kBlock), 0, st, (uint32_t)nw, claim, met,
                         sum);
      GX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_walk_final, dim3(gn), dim3(kBlock), 0, st, nn, E, nw, sum, urep, upos, isrep, ulen_of_rep,
                       ws + 1);
    GX_HIP(hipGetLastError());
    ph_end(b, PH_KWALK);
    uint32_t wr[4];
    GX_HIP(hipMemcpyAsync(wr, ws, sizeof(wr), hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    walked = wr[0] == 0 && wr[1] == 0;
    b->u_walkers = nw;
  }
  uint4* R[2] = {nullptr, nullptr};
  const unsigned jgrid = 8192;
  uint8_t* bflag = buf<uint8_t>(b->u_flag, jgrid);
  uint32_t* anyflag = buf<uint32_t>(b->u_any, 4);
  int rmax = 2;
  while ((1ull << rmax) < ne + 1) rmax++;
  int cur = 0, rounds = 0;
  if (!walked) {
  R[0] = buf<uint4>(b->u_ptr[0], ne + 1);
  R[1] = buf<uint4>(b->u_ptr[1], ne + 1);
  for (int attempt = 0; attempt < 2; attempt++) {
    cur = 0;
    if (nn) {
      hipLaunchKernelGGL(k_pj_init, dim3(grid_for(ne, kBlock, G * 4)), dim3(kBlock), 0, st, nn, lnk, msk, tand, cut,
                         R[0]);
      GX_HIP(hipGetLastError());
    }
    uint32_t act = nn ? 1 : 0;
    for (int r = 0; r <= rmax && act; r++) {
      unsigned g = grid_for(ne, kBlock, jgrid);
      pj_begin(b);
      hipLaunchKernelGGL(k_pj_round, dim3(g), dim3(kBlock), 0, st, ne, R[cur], R[cur ^ 1], bflag, errw + 2);
      GX_HIP(hipGetLastError());
      pj_end(b);
      hipLaunchKernelGGL(k_flag_or, dim3(1), dim3(kBlock), 0, st, bflag, g, anyflag);
      GX_HIP(hipGetLastError());
      cur ^= 1;
      rounds++;
      act = read1(st, anyflag);
    }
    if (!act) break;
    if (attempt == 1) throw GxError(GX_EDEVICE, "internal error: unitig chains did not terminate after cutting cycles");
    hipLaunchKernelGGL(k_pj_cut, dim3(grid_for(ne, kBlock, G)), dim3(kBlock), 0, st, ne, R[cur], lnk, cut);
    GX_HIP(hipGetLastError());
  }
  if (nn) {
    hipLaunchKernelGGL(k_pj_node_final, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, R[cur], urep, upos,
                       isrep, ulen_of_rep);
    GX_HIP(hipGetLastError());
  }
  }  // pointer jumping
  unsigned long long e = read1(st, errw);
  if (e != ~0ull) throw GxError(GX_EDEVICE, "internal error: neighbour k-mer missing from the table");
  b->u_succ_reads = read1(st, errw + 2);

  uint64_t* uid_of_rep = buf<uint64_t>(b->u_uid, nn + 1);
  exclusive_scan<uint32_t, uint64_t>(isrep, uid_of_rep, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
  uint64_t nu = read1(st, uid_of_rep + nn);
  uint32_t* urep_node = buf<uint32_t>(b->u_repnode, nu + 1);
  uint64_t* ulen = buf<uint64_t>(b->u_len, nu + 1);
  uint64_t* ubase = buf<uint64_t>(b->u_base, nu + 1);
  if (nn) {
    hipLaunchKernelGGL(k_rep_list, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, isrep, uid_of_rep,
                       ulen_of_rep, urep_node, ulen);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(ulen, ubase, nu, scratch_for(b, scan_scratch_bytes<uint64_t>(nu)), st);
  uint32_t* order = buf<uint32_t>(b->u_order, nn + 1);
  unsigned long long* ucov = buf<unsigned long long>(b->u_cov, nu + 1);
  uint32_t* uedge = buf<uint32_t>(b->u_edge, kUEdges * (nu + 1));
  uint8_t* uecnt = buf<uint8_t>(b->u_ecnt, nu + 1);
  if (nn) {
    hipLaunchKernelGGL(k_place, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, urep, upos, uid_of_rep, ubase,
                       order);
    GX_HIP(hipGetLastError());
  }
  // outer edges (latency-bound lookups) on the side stream, beside the
  // coverage sums and the heads; joined before the records
  side_init(b);
  stream_after(b->st2, st, b->ev_fork);
  if (nu) {
    hipLaunchKernelGGL(k_unitig_edges<W>, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, b->st2, nu, ubase, ulen,
                       order, klo, khi, msk, nix, urep, k, uedge, uecnt);
    GX_HIP(hipGetLastError());
  }
  if (nn) {
    uint32_t* ocov = buf<uint32_t>(b->u_isrep, nn + 1);  // isrep is read no more
    uint64_t* cpre = buf<uint64_t>(b->u_lenrep, nn + 1);  // nor ulen_of_rep
    hipLaunchKernelGGL(k_order_cov, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, order, cnt, ocov);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint32_t, uint64_t>(ocov, cpre, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
    hipLaunchKernelGGL(k_unitig_cov, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, nu, ubase, ulen, cpre, ucov);
    GX_HIP(hipGetLastError());
  }
  // heads
  uint64_t nh = b->nuheads;
  Key128* hk = buf<Key128>(b->u_hk, nh + 1);
  Key128* hk2 = buf<Key128>(b->u_hk2, nh + 1);
  uint32_t* hv = buf<uint32_t>(b->u_hv, nh + 1);
  uint32_t* hv2 = buf<uint32_t>(b->u_hv2, nh + 1);
  if (nn && nh) {
    hipLaunchKernelGGL(k_unitig_heads, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn,
                       b->hstart.as<uint64_t>(), b->hcu.as<uint32_t>(), b->hcf.as<uint32_t>(), b->uval.as<uint32_t>(),
                       b->has_hoff ? b->hoff.as<uint8_t>() : nullptr, urep, upos, uid_of_rep, b->rid.as<int64_t>(), k,
                       hk, hv, errw + 1);
    GX_HIP(hipGetLastError());
    const unsigned long long far = read1(st, errw + 1);
    if (far != ~0ull)
      throw GxError(GX_ENOTSUP, "a read head lies at letter " + std::to_string(far) +
                                    " of its unitig; read-head offsets hold 23 bits");
  }
  bool fl = radix_sort<Key128>(hk, hv, hk2, hv2, nh, 0, 12, scratch_for(b, radix_scratch_bytes<Key128>(nh)), st);
  if (fl) { std::swap(hk, hk2); std::swap(hv, hv2); }
  uint32_t* uflag = buf<uint32_t>(b->u_uflag, nh + 1);
  uint64_t* uposh = buf<uint64_t>(b->u_uposh, nh + 1);
  uint32_t* hu = buf<uint32_t>(b->u_hu, nu + 1);
  uint32_t* hf = buf<uint32_t>(b->u_hf, nu + 1);
  uint64_t* htot = buf<uint64_t>(b->u_htot, nu + 1);
  uint64_t* hstart = buf<uint64_t>(b->u_hstart, nu + 1);
  GX_HIP(hipMemsetAsync(hu, 0, (nu + 1) * 4, st));
  GX_HIP(hipMemsetAsync(hf, 0, (nu + 1) * 4, st));
  if (nh) {
    hipLaunchKernelGGL(k_uhead_unique, dim3(grid_for(nh, kBlock, G)), dim3(kBlock), 0, st, nh, hk, uflag, hu, hf);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint32_t, uint64_t>(uflag, uposh, nh, scratch_for(b, scan_scratch_bytes<uint64_t>(nh)), st);
  uint64_t* hkey_lo = buf<uint64_t>(b->u_hklo, nh + 1);
  uint32_t* hval_c = buf<uint32_t>(b->u_hvc, nh + 1);
  if (nh) {
    hipLaunchKernelGGL(k_uhead_compact, dim3(grid_for(nh, kBlock, G)), dim3(kBlock), 0, st, nh, uflag, uposh, hk, hv,
                       hkey_lo, hval_c);
    GX_HIP(hipGetLastError());
  }
  if (nu) {
    hipLaunchKernelGGL(k_head_total, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, nu, hu, hf, htot);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(htot, hstart, nu, scratch_for(b, scan_scratch_bytes<uint64_t>(nu)), st);
  // records (they list the outer edges)
  stream_after(st, b->st2, b->ev_join);
  USerCtx c;
  c.k = k; c.nb = nb; c.nu = nu; c.klo = klo; c.khi = khi; c.ubase = ubase; c.ulen = ulen; c.order = order;
  c.urep_node = urep_node; c.ucov = ucov; c.uedge = uedge; c.uecnt = uecnt; c.hstart = hstart; c.hu = hu; c.hf = hf;
  c.hkey = hkey_lo; c.hval = hval_c; c.mlen = b->mlen.as<uint32_t>(); c.wbase = b->wbase.as<uint64_t>();
  c.words = b->words.as<uint64_t>(); c.rid = b->rid.as<int64_t>();
  uint8_t* ends = buf<uint8_t>(b->u_ends, nn + 1);
  if (nn) {
    hipLaunchKernelGGL(k_node_ends<W>, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, klo, khi, k, ends);
    GX_HIP(hipGetLastError());
  }
  c.ends = ends;
  uint64_t* hbytes = buf<uint64_t>(b->u_hbytes, 2 * nu + 2);
  c.hbytes = hbytes;
  uint64_t* rsize = buf<uint64_t>(b->u_rsize, nu + 1);
  uint64_t* roff = buf<uint64_t>(b->u_roff, nu + 1);
  if (nu) {
    hipLaunchKernelGGL(k_uhead_bytes_long, dim3(grid_for(nu, kBlock / 64, 1u << 16)), dim3(kBlock), 0, st, c,
                       hbytes);
    GX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_useq_size, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, c, rsize);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(rsize, roff, nu, scratch_for(b, scan_scratch_bytes<uint64_t>(nu)), st);
  uint64_t total = read1(st, roff + nu);
  uint8_t* out = buf<uint8_t>(b->u_out, total + 1);
  if (nu) {
    uint64_t* seq_at = buf<uint64_t>(b->u_seqat, nu + 1);
    uint64_t* head_at = buf<uint64_t>(b->u_headat, 2 * nu + 2);
    // one block per kBlock unitigs (no grid-stride loop: every block stages its own span)
    hipLaunchKernelGGL(k_useq_write<W>, dim3(grid_for(nu, kBlock, 0xffffffffu)), dim3(kBlock), 0, st, c, roff, out,
                       seq_at, head_at);
    GX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_uheads_long, dim3(grid_for(nu, kBlock / 64, 1u << 16)), dim3(kBlock), 0, st, c, head_at,
                       out);
    GX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_useq_long<W>, dim3(grid_for(nu, kBlock / 64, 1u << 16)), dim3(kBlock), 0, st, c, seq_at, out);
    GX_HIP(hipGetLastError());
  }
  ph_end(b, PH_UNITIG);
  GX_HIP(hipStreamSynchronize(st));
  ph_collect(b);
  b->u_bytes = total;
  b->u_count = nu;
  b->u_rounds = (uint64_t)rounds;
}

template <class F>
static int guarded(gx_build* b, F&& f) {
  // a call that fails between a fork to the side stream and its join
  // leaves work on the side stream: it ends before the call returns, so no
  // later call reuses a buffer that it still reads
  struct SideDrain {
    gx_build* b;
    ~SideDrain() {
      if (b && b->st2) (void)hipStreamSynchronize(b->st2);
    }
  } drain{b};
  try {
    f();
    return GX_OK;
  } catch (const GxError& e) {
    if (b) b->err = e.what();
    return e.code;
  } catch (const HipError& e) {
    if (b) b->err = e.what();
    return e.code == hipErrorOutOfMemory ? GX_ENOMEM : GX_EDEVICE;
  } catch (const std::bad_alloc&) {
    if (b) b->err = "host allocation failed";
    return GX_ENOMEM;
  } catch (const std::exception& e) {
    if (b) b->err = e.what();
    return GX_EDEVICE;
  }
}

static void append_to(gx_build* b, DevBuf& dst, uint64_t& dlen, const void* src, size_t len, hipMemcpyKind kind,
                      size_t hint) {
  reclaim_forget(&dst);  // in use (a dead listing from an earlier build must not free it)
  size_t need = dlen + len + 64;
  if (need > dst.cap) {
    size_t ncap = std::max({need, hint, dst.cap * 3 / 2});
    void* np = nullptr;
    hipError_t e = hipMalloc(&np, ncap);
    while (e == hipErrorOutOfMemory && reclaim_dead(&dst, ncap)) {  // free the handle's dead buffers one by one
      (void)hipGetLastError();
      e = hipMalloc(&np, ncap);
    }
    if (e == hipErrorOutOfMemory && ncap > need) {  // the estimate or the growth step does not fit: what is needed
      (void)hipGetLastError();
      ncap = need;
      e = hipMalloc(&np, ncap);
    }
    if (tl_reclaim) tl_reclaim->allocs++;
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw HipError("hipMalloc of " + std::to_string(ncap) + " bytes: " + hipGetErrorString(e), (int)e);
    }
    if (dlen) GX_HIP(hipMemcpyAsync(np, dst.p, dlen, hipMemcpyDeviceToDevice, b->st));
    GX_HIP(hipStreamSynchronize(b->st));
    dst.release();
    dst.p = np;
    dst.cap = ncap;
  }
  if (len) GX_HIP(hipMemcpyAsync(dst.as<char>() + dlen, src, len, kind, b->st));
  GX_HIP(hipStreamSynchronize(b->st));
  dlen += len;
}

constexpr int kFmtFastqPair = 2;  // internal: two FASTQ files given through gx_build_add_fastq_pair

static void set_format(gx_build* b, int fmt) {
  if (b->finished) throw GxError(GX_ESTATE, "add_reads after finish; call gx_build_reset first");
  if (fmt != GX_FMT_READIDS && fmt != GX_FMT_FASTQ && fmt != kFmtFastqPair)
    throw GxError(GX_EINVAL, "unknown input format");
  if (b->fmt >= 0 && b->fmt != fmt)
    throw GxError(GX_EINVAL, "input format differs from earlier calls on this handle (one library per handle)");
  b->fmt = fmt;
}

static void append_text(gx_build* b, const void* src, size_t len, hipMemcpyKind kind, int fmt) {
  set_format(b, fmt);
  append_to(b, b->text, b->text_len, src, len, kind);
}

// Line ends of a text (same terminators as the readids parser).
static uint64_t line_ends(gx_build* b, const uint8_t* t, uint64_t n, DevBuf& termbuf, uint64_t** term_out) {
  hipStream_t st = b->st;
  unsigned lblocks = grid_for(n + 1, kLTile, 0xffffffffu);
  uint64_t* blk = buf<uint64_t>(b->blk, (size_t)lblocks + 1);
  uint64_t* blkb = buf<uint64_t>(b->blk2, (size_t)lblocks + 1);
  hipLaunchKernelGGL(k_line_count, dim3(lblocks), dim3(kBlock), 0, st, t, n, blk, nullptr);
  GX_HIP(hipGetLastError());
  exclusive_scan<uint64_t, uint64_t>(blk, blkb, lblocks, scratch_for(b, scan_scratch_bytes<uint64_t>(lblocks)), st);
  uint64_t nl = read1(st, blkb + lblocks);
  uint64_t* term = buf<uint64_t>(termbuf, nl + 1);
  hipLaunchKernelGGL(k_line_write, dim3(lblocks), dim3(kBlock), 0, st, t, n, blkb, term, nullptr);
  GX_HIP(hipGetLastError());
  *term_out = term;
  return nl;
}

// FASTQ (one file, or two line-aligned files) -> readids text in b->rtext.
static void fastq_convert(gx_build* b) {
  hipStream_t st = b->st;
  const bool paired = b->fmt == kFmtFastqPair;
  const uint8_t* t1 = b->text.as<uint8_t>();
  const uint8_t* t2 = nullptr;
  uint64_t *term1 = nullptr, *term2 = nullptr;
  uint64_t nl1 = line_ends(b, t1, b->text_len, b->fterm1, &term1);
  if (paired) {
    if (b->text2.cap == 0) b->text2.ensure(64);
    t2 = b->text2.as<uint8_t>();
    uint64_t nl2 = line_ends(b, t2, b->text2_len, b->fterm2, &term2);
    if (nl1 != nl2)
      throw GxError(GX_EINPUT, "Fastq files didn't have the same number of lines! (" + std::to_string(nl1) +
                                   " and " + std::to_string(nl2) + " lines)");
  }
  const uint64_t nrec = (nl1 + 2) / 4;  // lines 1, 5, 9, ...
  uint64_t* sz = buf<uint64_t>(b->fsz, nrec + 1);
  uint64_t* off = buf<uint64_t>(b->foff, nrec + 1);
  if (nrec) {
    hipLaunchKernelGGL(k_fq_size, dim3(grid_for(nrec, kBlock, 2048)), dim3(kBlock), 0, st, nrec, t1, term1, t2, term2,
                       sz);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(sz, off, nrec, scratch_for(b, scan_scratch_bytes<uint64_t>(nrec)), st);
  uint64_t total = read1(st, off + nrec);
  uint8_t* out = buf<uint8_t>(b->rtext, total + 64);
  if (nrec) {
    hipLaunchKernelGGL(k_fq_write, dim3(grid_for(nrec, kBlock / 64, 1u << 16)), dim3(kBlock), 0, st, nrec, t1, term1,
                       t2, term2, off, out);
    GX_HIP(hipGetLastError());
  }
  b->rtext_len = total;
}

// ===========================================================================
// Graph cleaning: REMOVE_BAD_COVERAGE, then MERGE (the pipeline step after
// the first merge, GenomixJobConf.java:511-517).  The reference removes each
// merged vertex whose average coverage lies below minCoverage or above
// maxCoverage, and its neighbours drop their edges to it
// (RemoveBadCoverageVertex.java:38-70); the next MERGE joins the rest.  On
// the device the node graph is cut instead: every node of a removed merge
// path goes, surviving nodes lose their edge bits towards removed nodes,
// and the unitig pass runs again over the surviving nodes (still in key
// order) -- the same merged records as merging the surviving vertices.
// ===========================================================================

// The removal rules, applied to the merged vertex (unitig) of each node.
enum CleanRule { CLEAN_COVERAGE = 0, CLEAN_TIP = 1 };
struct CleanArgs {
  int rule;
  float minc, maxc;  // CLEAN_COVERAGE: keep average coverage in [minc, maxc]
  uint64_t maxlen;   // CLEAN_TIP: remove tips of at most maxlen letters
  int k;
};

// keep[i]: node i's merged vertex survives the rule.  Coverage: the value
// its record carries.  Tip: one outer edge on one side, none on the other
// (Node.degree counts the FF+FR and RF+RR lists), and k + m - 1 letters.
__global__ void k_unitig_keep(uint64_t nn, const uint32_t* __restrict__ urep, const uint64_t* __restrict__ uid_of_rep,
                              const unsigned long long* __restrict__ ucov, const uint64_t* __restrict__ ulen,
                              const uint32_t* __restrict__ uedge, const uint8_t* __restrict__ uecnt, CleanArgs a,
                              uint32_t* __restrict__ keep) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t u = uid_of_rep[urep[i] & 0x7fffffffu];
    bool gone;
    if (a.rule == CLEAN_COVERAGE) {
      const float c = (float)((double)ucov[u] / (double)ulen[u]);
      gone = c < a.minc || c > a.maxc;
    } else {
      int deg[2] = {0, 0};
      for (int e = 0; e < uecnt[u]; e++) {
        const uint32_t t = uedge[kUEdges * u + e] >> kRepBits;
        deg[(t == ET_FF || t == ET_FR) ? 0 : 1]++;
      }
      const bool tip = (deg[0] == 1 && deg[1] == 0) || (deg[1] == 1 && deg[0] == 0);
      gone = tip && ulen[u] + (uint64_t)a.k - 1 <= a.maxlen;
    }
    keep[i] = gone ? 0u : 1u;
  }
}

// Read heads added to nodes by a bubble merge, grouped per node (unflipped
// ones first): per old node the counts and the first slot, per slot the
// mate index | flipped << 31 and the head's letter in the node's k-mer.
struct AddedHeads {
  const uint32_t *addu = nullptr, *addf = nullptr;
  const uint64_t* astart = nullptr;
  const uint32_t* aval = nullptr;
  const uint8_t* aoff = nullptr;
};

// Surviving nodes, compacted in key order, with the edge bits towards
// removed nodes cleared; added heads counted in.
template <int W>
__global__ void k_keep_nodes(uint64_t nn, const uint32_t* __restrict__ keep, const uint64_t* __restrict__ pos,
                             const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi,
                             const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ msk, NodeIndex ix, int k,
                             const uint32_t* __restrict__ hcu, const uint32_t* __restrict__ hcf, AddedHeads ah,
                             uint64_t* __restrict__ klo2, uint64_t* __restrict__ khi2, uint32_t* __restrict__ cnt2,
                             uint32_t* __restrict__ msk2, uint32_t* __restrict__ hcu2, uint32_t* __restrict__ hcf2,
                             uint64_t* __restrict__ htot2, unsigned long long* err) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!keep[i]) continue;
    const uint64_t j = pos[i];
    const KV v = node_key<W>(klo, khi, i);
    uint32_t m = msk[i];
    for (uint32_t t = 0; t < 4; t++)
      for (uint32_t x = 0; x < 4; x++) {
        if (!((m >> (4 * t + x)) & 1)) continue;
        const uint32_t q = node_lookup<W>(ix, neighbour<W>(v, k, t, x));
        if (q == kEnd) atomicMin(err, (unsigned long long)i);
        else if (!keep[q]) m &= ~(1u << (4 * t + x));
      }
    klo2[j] = klo[i];
    if (W == 2) khi2[j] = khi[i];
    cnt2[j] = cnt[i];
    msk2[j] = m;
    const uint32_t nu = hcu[i] + (ah.addu ? ah.addu[i] : 0u), nf = hcf[i] + (ah.addf ? ah.addf[i] : 0u);
    hcu2[j] = nu;
    hcf2[j] = nf;
    htot2[j] = (uint64_t)nu + nf;
  }
}

// Heads of surviving nodes: unflipped (kept, then added), flipped (kept,
// then added); with hoff2, each head's letter in its node's k-mer (hoff, or
// the build's 0 / k - 1 where the graph has none yet).
__global__ void k_keep_heads(uint64_t nn, const uint32_t* __restrict__ keep, const uint64_t* __restrict__ pos,
                             const uint64_t* __restrict__ hstart, const uint32_t* __restrict__ hcu,
                             const uint32_t* __restrict__ hcf, const uint32_t* __restrict__ uval,
                             const uint8_t* __restrict__ hoff, AddedHeads ah, int k,
                             const uint64_t* __restrict__ hstart2, uint32_t* __restrict__ uval2,
                             uint8_t* __restrict__ hoff2) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x) {
    if (!keep[i]) continue;
    const uint64_t a = hstart[i];
    uint64_t o = hstart2[pos[i]];
    const uint32_t au = ah.addu ? ah.addu[i] : 0u, af = ah.addf ? ah.addf[i] : 0u;
    const uint64_t a2 = ah.astart ? ah.astart[i] : 0;
    for (int f = 0; f < 2; f++) {
      const uint64_t s = a + (f ? hcu[i] : 0), n = f ? hcf[i] : hcu[i];
      for (uint64_t q = 0; q < n; q++, o++) {
        uval2[o] = uval[s + q];
        if (hoff2) hoff2[o] = hoff ? hoff[s + q] : (uint8_t)(f ? k - 1 : 0);
      }
      const uint64_t s2 = a2 + (f ? au : 0), n2 = f ? af : au;
      for (uint64_t q = 0; q < n2; q++, o++) {
        uval2[o] = ah.aval[s2 + q];
        hoff2[o] = ah.aoff[s2 + q];
      }
    }
  }
}

// The nodes with keep[i] set, with their heads and those of ah, become the
// handle's node graph (the build's own records are unchanged); then the
// node index and the unitig pass over it.
template <int W>
static void compact_graph(gx_build* b, uint32_t* keep, const AddedHeads& ah) {
  hipStream_t st = b->st;
  const uint64_t nn = b->nnodes;
  const unsigned G = b->grid;
  const NodeIndex nix{b->nix_dir.as<uint32_t>(), b->nix_slots.p, b->nb, 1u << (8 * b->nix_bkb), b->nix_bkb};
  unsigned long long* errw = buf<unsigned long long>(b->errw, 4);
  GX_HIP(hipMemsetAsync(errw, 0xff, 8, st));
  uint64_t* pos = buf<uint64_t>(b->c_pos, nn + 1);
  exclusive_scan<uint32_t, uint64_t>(keep, pos, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
  const uint64_t nn2 = read1(st, pos + nn);
  uint64_t* klo2 = buf<uint64_t>(b->c_klo, nn2 + 1);
  uint64_t* khi2 = W == 2 ? buf<uint64_t>(b->c_khi, nn2 + 1) : nullptr;
  uint32_t* cnt2 = buf<uint32_t>(b->c_cnt, nn2 + 1);
  uint32_t* msk2 = buf<uint32_t>(b->c_msk, nn2 + 1);
  uint32_t* hcu2 = buf<uint32_t>(b->c_hcu, nn2 + 1);
  uint32_t* hcf2 = buf<uint32_t>(b->c_hcf, nn2 + 1);
  uint64_t* htot2 = buf<uint64_t>(b->c_htot, nn2 + 1);
  uint64_t* hstart2 = buf<uint64_t>(b->c_hstart, nn2 + 1);
  if (nn) {
    hipLaunchKernelGGL(k_keep_nodes<W>, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, keep, pos,
                       b->klo.as<uint64_t>(), W == 2 ? b->khi.as<uint64_t>() : nullptr, b->cnt.as<uint32_t>(),
                       b->msk.as<uint32_t>(), nix, b->cfg.k, b->hcu.as<uint32_t>(), b->hcf.as<uint32_t>(), ah, klo2,
                       khi2, cnt2, msk2, hcu2, hcf2, htot2, errw);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(htot2, hstart2, nn2, scratch_for(b, scan_scratch_bytes<uint64_t>(nn2)), st);
  const uint64_t nh2 = read1(st, hstart2 + nn2);
  uint32_t* uval2 = buf<uint32_t>(b->c_uval, nh2 + 1);
  const bool offs = b->has_hoff || ah.aval;  // heads at letters other than the build's
  uint8_t* hoff2 = offs ? buf<uint8_t>(b->c_hoff, nh2 + 1) : nullptr;
  if (nn) {
    hipLaunchKernelGGL(k_keep_heads, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, keep, pos,
                       b->hstart.as<uint64_t>(), b->hcu.as<uint32_t>(), b->hcf.as<uint32_t>(), b->uval.as<uint32_t>(),
                       b->has_hoff ? b->hoff.as<uint8_t>() : nullptr, ah, b->cfg.k, hstart2, uval2, hoff2);
    GX_HIP(hipGetLastError());
  }
  if (read1(st, errw) != ~0ull) throw GxError(GX_EDEVICE, "internal error: neighbour k-mer missing from the graph");
  swap_buf(b->klo, b->c_klo);
  if (W == 2) swap_buf(b->khi, b->c_khi);
  swap_buf(b->cnt, b->c_cnt);
  swap_buf(b->msk, b->c_msk);
  swap_buf(b->hcu, b->c_hcu);
  swap_buf(b->hcf, b->c_hcf);
  swap_buf(b->hstart, b->c_hstart);
  swap_buf(b->uval, b->c_uval);
  if (offs) swap_buf(b->hoff, b->c_hoff);
  b->has_hoff = offs;
  b->nnodes = nn2;
  b->nuheads = nh2;
  build_node_index<W>(b, b->st, nn2, b->klo.as<uint64_t>(), W == 2 ? b->khi.as<uint64_t>() : nullptr, b->msk.as<uint32_t>(),
                      0, 256);
  run_unitigs<W>(b);
}

template <int W>
static void clean_graph(gx_build* b, const CleanArgs& ca) {
  hipStream_t st = b->st;
  const uint64_t nn = b->nnodes;
  uint32_t* keep = buf<uint32_t>(b->c_keep, nn + 1);
  if (nn) {
    hipLaunchKernelGGL(k_unitig_keep, dim3(grid_for(nn, kBlock, b->grid)), dim3(kBlock), 0, st, nn,
                       b->u_rep.as<uint32_t>(), b->u_uid.as<uint64_t>(), b->u_cov.as<unsigned long long>(),
                       b->u_len.as<uint64_t>(), b->u_edge.as<uint32_t>(), b->u_ecnt.as<uint8_t>(), ca, keep);
    GX_HIP(hipGetLastError());
  }
  compact_graph<W>(b, keep, AddedHeads{});
}

// ===========================================================================
// BUBBLE (SimpleBubbleMergeVertex.java:51-320) over the merged graph of the
// last unitig pass, applied to the node graph: removed bubbles drop their
// nodes, the kept one (the top) takes their k-mer counts on its first node
// and their read heads, scaled to its length, as letters of its nodes.
// ===========================================================================

constexpr uint32_t kBubMaxLetters = 4096;  // longest bubble sequence (edit-distance row in LDS)

__device__ inline int et_side_d(uint32_t t) { return (t == ET_FF || t == ET_FR) ? 0 : 1; }
__device__ inline uint32_t et_mirror_d(uint32_t t) { return t == ET_FF ? ET_RR : t == ET_RR ? ET_FF : t; }
__device__ inline bool et_flips_d(uint32_t t) { return t == ET_FR || t == ET_RF; }  // EDGETYPE.causesFlip

// Bubbles (detectBubble :245-252, sendBubbleAndMajorVertexMsgToMinorVertex
// :51-79): one outer edge on each side, to two different merged vertices,
// and fewer than maxlen letters.  The neighbour whose key is not smaller
// (VKmer.compareTo, byte order) is the major vertex.  bmin/bmaj: minor /
// major unitig (kEnd: not a bubble); btyp: majorToBubble | minorToBubble << 2.
template <int W>
__global__ void k_bub_cand(uint64_t nu, const uint64_t* __restrict__ ulen, const uint32_t* __restrict__ uedge,
                           const uint8_t* __restrict__ uecnt, const uint64_t* __restrict__ uid_of_rep,
                           const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi, int k,
                           uint64_t maxlen, uint32_t* __restrict__ bmin, uint32_t* __restrict__ bmaj,
                           uint8_t* __restrict__ btyp) {
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < nu; u += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t ent[2] = {kEnd, kEnd};
    int deg[2] = {0, 0};
    for (int e = 0; e < uecnt[u]; e++) {
      const uint32_t x = uedge[kUEdges * u + e];
      const int sd = et_side_d(x >> kRepBits);
      if (deg[sd]++ == 0) ent[sd] = x;
    }
    uint32_t mn = kEnd, mj = kEnd;
    uint8_t ty = 0;
    if (deg[0] == 1 && deg[1] == 1 && ulen[u] + (uint64_t)k - 1 < maxlen) {
      const uint32_t rf = ent[0] & kRepMask, rr = ent[1] & kRepMask;
      if (rf != rr) {
        const bool fwd_major = kle<W>(node_key<W>(klo, khi, rr), node_key<W>(klo, khi, rf));
        const uint32_t emaj = fwd_major ? ent[0] : ent[1], emin = fwd_major ? ent[1] : ent[0];
        mj = (uint32_t)uid_of_rep[emaj & kRepMask];
        mn = (uint32_t)uid_of_rep[emin & kRepMask];
        ty = (uint8_t)(et_mirror_d(emaj >> kRepBits) | (et_mirror_d(emin >> kRepBits) << 2));
      }
    }
    bmin[u] = mn;
    bmaj[u] = mj;
    btyp[u] = ty;
  }
}

struct BubCtx {
  const uint64_t *ulen, *uid_of_rep;
  const unsigned long long* ucov;
  const uint32_t *uedge, *urep_node, *bmin, *bmaj;
  const uint8_t *uecnt, *btyp;
};

__device__ inline float bub_cov(const BubCtx& c, uint32_t w) {
  return (float)((double)c.ucov[w] / (double)c.ulen[w]);
}
// isValidMajorAndMinor (:98-105): same side of the major and of the minor.
__device__ inline uint32_t bub_class(const BubCtx& c, uint32_t w) {
  return (uint32_t)et_side_d(c.btyp[w] & 3u) | ((uint32_t)et_side_d((c.btyp[w] >> 2) & 3u) << 1);
}

// The bubbles whose minor vertex is v, in processing order: by major
// vertex, then by class (bubbles of different classes never meet in
// processSimilarSet), then by decreasing coverage (SortByCoverage, stable;
// before it the minor's message order, taken as the merged records' order,
// i.e. node order).
__device__ inline uint32_t bub_set(const BubCtx& c, uint64_t v, uint32_t* w) {
  uint32_t n = 0;
  for (int e = 0; e < c.uecnt[v]; e++) {
    const uint32_t x = (uint32_t)c.uid_of_rep[c.uedge[kUEdges * v + e] & kRepMask];
    if (c.bmin[x] != (uint32_t)v) continue;
    bool dup = false;
    for (uint32_t a = 0; a < n; a++) dup |= w[a] == x;
    if (dup) continue;
    const uint32_t cx = bub_class(c, x);
    const float fx = bub_cov(c, x);
    uint32_t p = n++;
    while (p > 0) {
      const uint32_t y = w[p - 1];
      const uint32_t cy = bub_class(c, y);
      const float fy = bub_cov(c, y);
      const bool before = c.bmaj[y] != c.bmaj[x] ? c.bmaj[y] < c.bmaj[x]
                          : cy != cx             ? cy < cx
                          : fy != fx             ? fy > fx
                                                 : c.urep_node[y] < c.urep_node[x];
      if (before) break;
      w[p] = y;
      p--;
    }
    w[p] = x;
  }
  return n;
}

__device__ inline bool bub_same_group(const BubCtx& c, uint32_t a, uint32_t b) {
  return c.bmaj[a] == c.bmaj[b] && bub_class(c, a) == bub_class(c, b);
}

// Pairs to compare, per minor vertex: every pair within a group of the same
// major and class.  pbase null: count them; else write (first, second,
// opposite orientations).
__global__ void k_bub_pairs(uint64_t nu, BubCtx c, uint32_t* __restrict__ npairs,
                            const uint64_t* __restrict__ pbase, uint4* __restrict__ pairs) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nu; v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t w[kUEdges];
    const uint32_t n = bub_set(c, v, w);
    uint64_t q = pbase ? pbase[v] : 0;
    uint32_t cntp = 0;
    for (uint32_t i = 0; i < n; i++)
      for (uint32_t j = i + 1; j < n && bub_same_group(c, w[i], w[j]); j++) {
        if (pbase) {
          const bool flip = et_flips_d(c.btyp[w[i]] & 3u) != et_flips_d(c.btyp[w[j]] & 3u);
          pairs[q++] = make_uint4(w[i], w[j], flip ? 1u : 0u, 0u);
        }
        cntp++;
      }
    if (!pbase) npairs[v] = cntp;
  }
}

// Letter j of unitig u's sequence (its first k-mer, then the last letter of
// each next one).
template <int W>
__device__ inline uint32_t unitig_letter(const uint32_t* order, const uint64_t* ubase, const uint64_t* klo,
                                         const uint64_t* khi, int k, uint32_t u, uint32_t j) {
  const uint32_t p = j < (uint32_t)k ? 0u : j - (uint32_t)k + 1;
  const uint32_t e = order[ubase[u] + p];
  return oriented_letter<W>(klo, khi, e & 0x7fffffffu, !(e & 0x80000000u), k, (int)(j - p));
}

// VKmer.editDistance (VKmer.java:711-731) of each pair, against the
// reverse complement of the second when the two are stored in opposite
// orientations (VKmer.fracDissimilar :919-928).  A wave per pair, a row of
// the distance matrix at a time: D[r][c] = min over c' <= c of
// T[c'] + (c - c'), T[c] = min(D[r-1][c] + 1, D[r-1][c-1] + mismatch),
// i.e. a prefix minimum of T[c] - c across the lanes.
template <int W>
__global__ void __launch_bounds__(64) k_bub_dist(uint64_t np, const uint4* __restrict__ pairs,
                                                 const uint32_t* __restrict__ order, const uint64_t* __restrict__ ubase,
                                                 const uint64_t* __restrict__ ulen, const uint64_t* __restrict__ klo,
                                                 const uint64_t* __restrict__ khi, int k, uint32_t* __restrict__ ed) {
  __shared__ uint16_t row[kBubMaxLetters + 1];
  __shared__ uint8_t sa[kBubMaxLetters], sb[kBubMaxLetters];
  const int lane = threadIdx.x;
  constexpr int kBig = 1 << 28;
  for (uint64_t p = blockIdx.x; p < np; p += gridDim.x) {
    const uint4 pr = pairs[p];
    const uint32_t La = (uint32_t)(ulen[pr.x] + k - 1), Lb = (uint32_t)(ulen[pr.y] + k - 1);
    for (uint32_t j = lane; j < La; j += 64) sa[j] = (uint8_t)unitig_letter<W>(order, ubase, klo, khi, k, pr.x, j);
    for (uint32_t j = lane; j < Lb; j += 64)
      sb[j] = (uint8_t)(pr.z ? 3u - unitig_letter<W>(order, ubase, klo, khi, k, pr.y, Lb - 1 - j)
                             : unitig_letter<W>(order, ubase, klo, khi, k, pr.y, j));
    for (uint32_t c = lane; c <= Lb; c += 64) row[c] = (uint16_t)c;
    __syncthreads();
    for (uint32_t r = 1; r <= La; r++) {
      const uint32_t ar = sa[r - 1];
      int carry_old = 0, carry_min = kBig;
      for (uint32_t c0 = 0; c0 <= Lb; c0 += 64) {
        const uint32_t c = c0 + lane;
        const bool valid = c <= Lb;
        const int old = valid ? (int)row[c] : kBig;
        int oldm1 = __shfl_up(old, 1);
        if (lane == 0) oldm1 = carry_old;
        int v = kBig;
        if (valid) {
          const int t = c == 0 ? (int)r : min(old + 1, oldm1 + (ar != sb[c - 1] ? 1 : 0));
          v = t - (int)c;
        }
        for (int d = 1; d < 64; d <<= 1) {
          const int o = __shfl_up(v, d);
          if (lane >= d) v = min(v, o);
        }
        v = min(v, carry_min);
        if (valid) row[c] = (uint16_t)(v + (int)c);
        carry_old = __shfl(old, 63);
        carry_min = __shfl(v, 63);
      }
    }
    __syncthreads();
    if (lane == 0) ed[p] = row[Lb];
    __syncthreads();
  }
}

// processSimilarSet (:163-240) per minor vertex: in each group the first
// bubble left is the top; every later one whose edit distance over the
// shorter length is at most maxd is removed into it (into = top | opposite
// orientations << 31); the top leaves the group.
__global__ void k_bub_decide(uint64_t nu, BubCtx c, const uint64_t* __restrict__ pbase,
                             const uint32_t* __restrict__ ed, int k, float maxd, uint32_t* __restrict__ into) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nu; v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t w[kUEdges];
    const uint32_t n = bub_set(c, v, w);
    uint64_t q = pbase[v];
    uint32_t alive = (1u << n) - 1u;
    for (uint32_t i = 0; i < n; i++) {
      uint64_t qi = q;  // pair (i, i+1) of this group
      uint32_t j = i + 1;
      for (; j < n && bub_same_group(c, w[i], w[j]); j++, qi++) {
        if (!((alive >> i) & 1u) || !((alive >> j) & 1u)) continue;
        const uint32_t li = (uint32_t)(c.ulen[w[i]] + k - 1), lj = (uint32_t)(c.ulen[w[j]] + k - 1);
        const float frac = __fdiv_rn((float)ed[qi], (float)min(li, lj));
        if (frac <= maxd) {
          const bool flip = et_flips_d(c.btyp[w[i]] & 3u) != et_flips_d(c.btyp[w[j]] & 3u);
          into[w[j]] = w[i] | (flip ? 0x80000000u : 0u);
          alive &= ~(1u << j);
        }
      }
      q = qi;
      alive &= ~(1u << i);
    }
  }
}

// Nodes of removed bubbles are dropped; the top's first node takes their
// k-mer counts (Node.addCoverage :290-294 in counts); read heads per
// removed bubble counted.
__global__ void k_bub_apply(uint64_t nn, uint64_t nu, const uint32_t* __restrict__ urep,
                            const uint64_t* __restrict__ uid_of_rep, const uint32_t* __restrict__ into,
                            const uint32_t* __restrict__ order, const uint64_t* __restrict__ ubase,
                            const unsigned long long* __restrict__ ucov, const uint32_t* __restrict__ hu,
                            const uint32_t* __restrict__ hf, uint32_t* __restrict__ keep, uint32_t* __restrict__ cnt,
                            uint32_t* __restrict__ hn) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn || i < nu;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (i < nn) keep[i] = into[uid_of_rep[urep[i] & 0x7fffffffu]] == kEnd ? 1u : 0u;
    if (i < nu) {
      const uint32_t t = into[i];
      hn[i] = t == kEnd ? 0u : hu[i] + hf[i];
      if (t != kEnd) atomicAdd(&cnt[order[ubase[t & 0x7fffffffu]] & 0x7fffffffu], (uint32_t)ucov[i]);
    }
  }
}

// ReadHeadSet.unionUpdate(other, lengthFactor, flip, otherLength)
// (:80-101) for the heads of each removed bubble: offset o (counted from
// the other end when the two are stored in opposite orientations, the set
// swapped) becomes (int)((o + 1) * f - f) in float, f = top letters /
// bubble letters; the head goes to the top's node covering that letter
// (the last node past its end), as the letter's place in the node's k-mer.
__global__ void k_bub_heads(uint64_t nu, const uint32_t* __restrict__ into, const uint64_t* __restrict__ hb,
                            const uint64_t* __restrict__ hstart, const uint32_t* __restrict__ hu,
                            const uint32_t* __restrict__ hf, const uint64_t* __restrict__ hkey,
                            const uint32_t* __restrict__ hval, const uint64_t* __restrict__ ulen,
                            const uint32_t* __restrict__ order, const uint64_t* __restrict__ ubase, int k,
                            uint32_t* __restrict__ hnode, uint32_t* __restrict__ hmv, uint8_t* __restrict__ hofs,
                            uint32_t* __restrict__ addu, uint32_t* __restrict__ addf) {
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nu; w += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t x = into[w];
    if (x == kEnd) continue;
    const uint32_t t = x & 0x7fffffffu;
    const bool flip = (x >> 31) != 0;
    const uint32_t Lo = (uint32_t)(ulen[w] + k - 1), Lt = (uint32_t)(ulen[t] + k - 1);
    const uint64_t mt = ulen[t];
    const float fac = __fdiv_rn((float)Lt, (float)Lo);
    const uint32_t n = hu[w] + hf[w];
    for (uint32_t j = 0; j < n; j++) {
      const uint64_t key = hkey[hstart[w] + j];
      const uint32_t off = (uint32_t)((key >> 30) & ((1ull << 23) - 1));
      const bool nf = ((key >> 53) & 1) != 0;
      const uint32_t o = flip ? Lo - 1 - off : off;
      const float y = __fsub_rn(__fmul_rn((float)(o + 1), fac), fac);
      const uint64_t a = (uint64_t)(int)y;  // a letter of the top's sequence
      const bool nf2 = flip ? !nf : nf;
      const uint64_t q = a < mt - 1 ? a : mt - 1;
      const uint32_t e = order[ubase[t] + q];
      const bool fwd = !(e & 0x80000000u);
      const uint32_t rel = (uint32_t)(a - q);
      const bool fn = fwd ? nf2 : !nf2;
      const uint64_t s = hb[w] + j;
      hnode[s] = e & 0x7fffffffu;
      hmv[s] = (hval[hstart[w] + j] & 0x7fffffffu) | (fn ? 0x80000000u : 0u);
      hofs[s] = (uint8_t)(fwd ? rel : (uint32_t)(k - 1) - rel);
      atomicAdd(fn ? &addf[e & 0x7fffffffu] : &addu[e & 0x7fffffffu], 1u);
    }
  }
}

__global__ void k_bub_total(uint64_t nn, const uint32_t* __restrict__ addu, const uint32_t* __restrict__ addf,
                            uint32_t* __restrict__ tot) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (uint64_t)gridDim.x * blockDim.x)
    tot[i] = addu[i] + addf[i];
}

// Added heads into their nodes' slots (unflipped first).
__global__ void k_bub_place(uint64_t na, const uint32_t* __restrict__ hnode, const uint32_t* __restrict__ hmv,
                            const uint8_t* __restrict__ hofs, const uint64_t* __restrict__ astart,
                            const uint32_t* __restrict__ addu, uint32_t* __restrict__ cur,
                            uint32_t* __restrict__ aval, uint8_t* __restrict__ aoff) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < na; h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = hnode[h];
    const bool f = (hmv[h] >> 31) != 0;
    const uint64_t s = astart[i] + (f ? addu[i] : 0u) + atomicAdd(&cur[2 * (uint64_t)i + (f ? 1 : 0)], 1u);
    aval[s] = hmv[h];
    aoff[s] = hofs[h];
  }
}

template <int W>
static void bubble_graph(gx_build* b, float maxd, uint64_t maxlen) {
  hipStream_t st = b->st;
  const uint64_t nn = b->nnodes, nu = b->u_count;
  const unsigned G = b->grid;
  const int k = b->cfg.k;
  uint32_t* bmin = buf<uint32_t>(b->bb_min, nu + 1);
  uint32_t* bmaj = buf<uint32_t>(b->bb_maj, nu + 1);
  uint8_t* btyp = buf<uint8_t>(b->bb_typ, nu + 1);
  uint32_t* np = buf<uint32_t>(b->bb_np, nu + 1);
  uint64_t* pbase = buf<uint64_t>(b->bb_pbase, nu + 1);
  uint32_t* into = buf<uint32_t>(b->bb_into, nu + 1);
  BubCtx c;
  c.ulen = b->u_len.as<uint64_t>();
  c.uid_of_rep = b->u_uid.as<uint64_t>();
  c.ucov = b->u_cov.as<unsigned long long>();
  c.uedge = b->u_edge.as<uint32_t>();
  c.urep_node = b->u_repnode.as<uint32_t>();
  c.uecnt = b->u_ecnt.as<uint8_t>();
  c.bmin = bmin;
  c.bmaj = bmaj;
  c.btyp = btyp;
  const uint32_t* order = b->u_order.as<uint32_t>();
  const uint64_t* ubase = b->u_base.as<uint64_t>();
  const uint64_t* klo = b->klo.as<uint64_t>();
  const uint64_t* khi = W == 2 ? b->khi.as<uint64_t>() : nullptr;
  uint64_t npairs = 0;
  if (nu) {
    hipLaunchKernelGGL(k_bub_cand<W>, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, nu, c.ulen, c.uedge,
                       c.uecnt, c.uid_of_rep, klo, khi, k, maxlen, bmin, bmaj, btyp);
    GX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bub_pairs, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, nu, c, np, nullptr, nullptr);
    GX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, into, 0, nu, kEnd);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint32_t, uint64_t>(np, pbase, nu, scratch_for(b, scan_scratch_bytes<uint64_t>(nu)), st);
  npairs = read1(st, pbase + nu);
  b->bub_pairs = npairs;
  if (npairs) {
    uint4* pairs = buf<uint4>(b->bb_pairs, npairs);
    uint32_t* ed = buf<uint32_t>(b->bb_ed, npairs);
    hipLaunchKernelGGL(k_bub_pairs, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, nu, c, np, pbase, pairs);
    GX_HIP(hipGetLastError());
    const unsigned gd = (unsigned)std::min<uint64_t>(npairs, 1u << 16);
    hipLaunchKernelGGL(k_bub_dist<W>, dim3(gd), dim3(64), 0, st, npairs, pairs, order, ubase, c.ulen, klo, khi, k, ed);
    GX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bub_decide, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, nu, c, pbase, ed, k, maxd,
                       into);
    GX_HIP(hipGetLastError());
  }
  uint32_t* keep = buf<uint32_t>(b->c_keep, nn + 1);
  uint32_t* hn = buf<uint32_t>(b->bb_hn, nu + 1);
  uint64_t* hb = buf<uint64_t>(b->bb_hb, nu + 1);
  if (nn || nu) {
    hipLaunchKernelGGL(k_bub_apply, dim3(grid_for(std::max(nn, nu), kBlock, G)), dim3(kBlock), 0, st, nn, nu,
                       b->u_rep.as<uint32_t>(), c.uid_of_rep, into, order, ubase, c.ucov, b->u_hu.as<uint32_t>(),
                       b->u_hf.as<uint32_t>(), keep, b->cnt.as<uint32_t>(), hn);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint32_t, uint64_t>(hn, hb, nu, scratch_for(b, scan_scratch_bytes<uint64_t>(nu)), st);
  const uint64_t na = read1(st, hb + nu);
  AddedHeads ah;
  if (na) {
    uint32_t* addu = buf<uint32_t>(b->bb_addu, nn + 1);
    uint32_t* addf = buf<uint32_t>(b->bb_addf, nn + 1);
    uint32_t* cur = buf<uint32_t>(b->bb_cur, 2 * nn + 2);
    uint32_t* tot = buf<uint32_t>(b->bb_tot, nn + 1);
    uint64_t* astart = buf<uint64_t>(b->bb_astart, nn + 1);
    uint32_t* hnode = buf<uint32_t>(b->bb_hnode, na);
    uint32_t* hmv = buf<uint32_t>(b->bb_hmv, na);
    uint8_t* hofs = buf<uint8_t>(b->bb_hofs, na);
    uint32_t* aval = buf<uint32_t>(b->bb_aval, na);
    uint8_t* aoff = buf<uint8_t>(b->bb_aoff, na);
    GX_HIP(hipMemsetAsync(addu, 0, (nn + 1) * 4, st));
    GX_HIP(hipMemsetAsync(addf, 0, (nn + 1) * 4, st));
    GX_HIP(hipMemsetAsync(cur, 0, (2 * nn + 2) * 4, st));
    hipLaunchKernelGGL(k_bub_heads, dim3(grid_for(nu, kBlock, G)), dim3(kBlock), 0, st, nu, into, hb,
                       b->u_hstart.as<uint64_t>(), b->u_hu.as<uint32_t>(), b->u_hf.as<uint32_t>(),
                       b->u_hklo.as<uint64_t>(), b->u_hvc.as<uint32_t>(), c.ulen, order, ubase, k, hnode, hmv, hofs,
                       addu, addf);
    GX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bub_total, dim3(grid_for(nn, kBlock, G)), dim3(kBlock), 0, st, nn, addu, addf, tot);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint32_t, uint64_t>(tot, astart, nn, scratch_for(b, scan_scratch_bytes<uint64_t>(nn)), st);
    hipLaunchKernelGGL(k_bub_place, dim3(grid_for(na, kBlock, G)), dim3(kBlock), 0, st, na, hnode, hmv, hofs, astart,
                       addu, cur, aval, aoff);
    GX_HIP(hipGetLastError());
    ah.addu = addu;
    ah.addf = addf;
    ah.astart = astart;
    ah.aval = aval;
    ah.aoff = aoff;
  }
  compact_graph<W>(b, keep, ah);
}

// ===========================================================================
// Graph statistics and FASTA dump (GraphStatistics.java:79-146,
// DriverUtils.dumpGraph :85-128) over the node graph or the last merge.
// A vertex's properties: out-degree (FF + FR entries), in-degree (RF + RR),
// sequence length (the internal k-mer, or k), average coverage, unflipped
// and flipped read heads, self edges per type.
// ===========================================================================

struct VertexCtx {
  int merged;  // 0: nodes, 1: unitigs of the last merge
  int k;
  uint64_t n;
  const uint64_t *klo, *khi;
  // nodes
  const uint32_t *cnt, *msk, *hcu, *hcf;
  // unitigs
  const uint64_t* ulen;
  const unsigned long long* ucov;
  const uint32_t* uedge;
  const uint8_t* uecnt;
  const uint32_t *urep_node, *hu, *hf;
  const uint64_t* ubase;
  const uint32_t* order;
};

struct VProps {
  uint32_t deg[2];  // [0] forward (out), [1] reverse (in)
  uint64_t len;
  float cov;
  uint32_t nu, nf;
  uint32_t self[4];
};

template <int W, bool SELF>
__device__ inline VProps vertex_props(const VertexCtx& c, uint64_t i) {
  VProps p{};
  if (!c.merged) {
    const uint32_t m = c.msk[i];
    p.deg[0] = __popc(m & 0xFFu);
    p.deg[1] = __popc((m >> 8) & 0xFFu);
    p.len = (uint64_t)c.k;
    p.cov = (float)(c.cnt[i] > (1u << 24) ? (1u << 24) : c.cnt[i]);
    p.nu = c.hcu[i];
    p.nf = c.hcf[i];
    if (SELF) {
      const KV v = node_key<W>(c.klo, c.khi, i);
      for (uint32_t t = 0; t < 4; t++)
        for (uint32_t x = 0; x < 4; x++)
          if (((m >> (4 * t + x)) & 1) && keq<W>(neighbour<W>(v, c.k, t, x), v)) p.self[t]++;
    }
  } else {
    const uint32_t rep = c.urep_node[i];
    for (int e = 0; e < c.uecnt[i]; e++) {
      const uint32_t x = c.uedge[kUEdges * i + e], t = x >> kRepBits;
      p.deg[(t == ET_FF || t == ET_FR) ? 0 : 1]++;
      if ((x & kRepMask) == rep) p.self[t]++;
    }
    p.len = c.ulen[i] + (uint64_t)c.k - 1;
    p.cov = (float)((double)c.ucov[i] / (double)c.ulen[i]);
    p.nu = c.hu[i];
    p.nf = c.hf[i];
  }
  return p;
}

// Java Math.round(float) (floor of x + 0.5f, in float) and the seed score
// (Node.calculateSeedScore, Node.java:859-862: length x read heads).
__device__ inline int64_t java_round(float x) { return (int64_t)floorf(x + 0.5f); }

enum { ST_DEGREE, ST_KLEN, ST_COV, ST_UNFLIPPED, ST_FLIPPED, ST_SEED, ST_KLEN_F, ST_COV_F, ST_SEED_F, ST_KLEN_R,
       ST_COV_R, ST_SEED_R, NSTAT };
enum { TOT_NODES, TOT_PATH, TOT_TIPS_F, TOT_TIPS_R, TOT_TIPS_BOTH, TOT_TIPS_ONE, TOT_SELF_FF, NTOT = TOT_SELF_FF + 4 };

// The value of statistic s for vertex i, or nothing (GraphStatistics
// counts scaffoldSeedScore only for coverage within mean 0 +- 1000, and the
// -with-d statistics only where degree(d) != 0).
template <int W>
__device__ inline bool stat_value(const VertexCtx& c, const VProps& p, int s, int64_t* out) {
  const bool win = p.cov >= -1000.f && p.cov <= 1000.f;
  const int64_t seed = (int64_t)(int32_t)((uint32_t)p.len * (p.nu + p.nf));
  switch (s) {
    case ST_DEGREE: *out = p.deg[0] + p.deg[1]; return true;
    case ST_KLEN: *out = (int64_t)p.len; return true;
    case ST_COV: *out = java_round(p.cov); return true;
    case ST_UNFLIPPED: *out = p.nu; return true;
    case ST_FLIPPED: *out = p.nf; return true;
    case ST_SEED: *out = seed; return win;
    case ST_KLEN_F: *out = (int64_t)p.len; return p.deg[0] != 0;
    case ST_COV_F: *out = java_round(p.cov); return p.deg[0] != 0;
    case ST_SEED_F: *out = seed; return p.deg[0] != 0 && win;
    case ST_KLEN_R: *out = (int64_t)p.len; return p.deg[1] != 0;
    case ST_COV_R: *out = java_round(p.cov); return p.deg[1] != 0;
    default: *out = seed; return p.deg[1] != 0 && win;
  }
}

// Values of statistic s (as value + 2^63, ordered like the signed values)
// and whether each vertex has one; with their count, sum, least and
// largest value.
// Wave sums, minima and maxima of 64-bit values (lane 0 holds the result).
__device__ inline uint64_t wave_sum(uint64_t x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ inline uint64_t wave_min(uint64_t x) {
  for (int o = 32; o > 0; o >>= 1) x = min(x, (uint64_t)__shfl_xor(x, o, 64));
  return x;
}
__device__ inline uint64_t wave_max(uint64_t x) {
  for (int o = 32; o > 0; o >>= 1) x = max(x, (uint64_t)__shfl_xor(x, o, 64));
  return x;
}

template <int W>
__global__ void k_stat_values(VertexCtx c, int s, uint64_t* __restrict__ keys, uint32_t* __restrict__ has,
                              unsigned long long* __restrict__ acc) {
  uint64_t cnt = 0, sum = 0, lo = ~0ull, hi = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += (uint64_t)gridDim.x * blockDim.x) {
    const VProps p = vertex_props<W, false>(c, i);
    int64_t v;
    const bool in = stat_value<W>(c, p, s, &v);
    const uint64_t key = (uint64_t)v + 0x8000000000000000ull;
    keys[i] = key;
    has[i] = in ? 1u : 0u;
    if (in) {
      cnt++;
      sum += (uint64_t)v;
      lo = min(lo, key);
      hi = max(hi, key);
    }
  }
  // one set of atomics per wave
  cnt = wave_sum(cnt);
  sum = wave_sum(sum);
  lo = wave_min(lo);
  hi = wave_max(hi);
  if ((threadIdx.x & 63) == 0 && cnt) {
    atomicAdd(&acc[0], (unsigned long long)cnt);
    atomicAdd(&acc[1], (unsigned long long)sum);
    atomicMin(&acc[3], (unsigned long long)lo);
    atomicMax(&acc[2], (unsigned long long)hi);
  }
}

// The values present, compacted and offset by the least: sortable on the
// few low bytes their range needs.
__global__ void k_stat_compact(uint64_t n, const uint64_t* __restrict__ keys, const uint32_t* __restrict__ has,
                               const uint64_t* __restrict__ pos, uint64_t lo, uint64_t* __restrict__ out,
                               uint32_t* __restrict__ vals) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    if (has[i]) {
      out[pos[i]] = keys[i] - lo;
      vals[pos[i]] = 0;
    }
}

template <int W>
__global__ void k_stat_totals(VertexCtx c, unsigned long long* __restrict__ tot) {
  uint64_t t[NTOT] = {};
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += (uint64_t)gridDim.x * blockDim.x) {
    const VProps p = vertex_props<W, true>(c, i);
    t[TOT_NODES]++;
    t[TOT_PATH] += p.deg[0] == 1 && p.deg[1] == 1;
    t[TOT_TIPS_F] += p.deg[0] == 0;
    t[TOT_TIPS_R] += p.deg[1] == 0;
    t[TOT_TIPS_BOTH] += p.deg[0] == 0 && p.deg[1] == 0;
    t[TOT_TIPS_ONE] += (p.deg[0] == 0) != (p.deg[1] == 0);
    for (int e = 0; e < 4; e++) t[TOT_SELF_FF + e] += p.self[e];
  }
#pragma unroll
  for (int q = 0; q < NTOT; q++) {
    const uint64_t x = wave_sum(t[q]);
    if ((threadIdx.x & 63) == 0 && x) atomicAdd(&tot[q], (unsigned long long)x);
  }
}

// Distinct sorted values and their counts (first `m` keys are the values).
__global__ void k_bin_starts(uint64_t m, const uint64_t* __restrict__ keys, uint32_t* __restrict__ flag) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_bin_write(uint64_t m, const uint64_t* __restrict__ keys, uint64_t lo, const uint32_t* __restrict__ flag,
                            const uint64_t* __restrict__ pos, int64_t* __restrict__ bval, uint64_t* __restrict__ bstart) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
    if (flag[i]) {
      bval[pos[i]] = (int64_t)(keys[i] + lo - 0x8000000000000000ull);
      bstart[pos[i]] = i;
    }
}

// FASTA: ">node_" + key letters + "\n" + sequence letters + "\n" per vertex
// (the internal k-mer of a merged vertex, else its key).
constexpr char kAcgt[4] = {'A', 'C', 'G', 'T'};

template <int W>
__device__ inline uint64_t fasta_seq_len(const VertexCtx& c, uint64_t i) {
  return c.merged ? c.ulen[i] + (uint64_t)c.k - 1 : (uint64_t)c.k;
}

template <int W>
__global__ void k_fasta_size(VertexCtx c, uint64_t* __restrict__ sz) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < c.n; i += (uint64_t)gridDim.x * blockDim.x)
    sz[i] = 6 + (uint64_t)c.k + 1 + fasta_seq_len<W>(c, i) + 1;
}

// Letter j of vertex i's sequence.
template <int W>
__device__ inline uint32_t fasta_letter(const VertexCtx& c, uint64_t i, uint64_t j) {
  if (!c.merged) return oriented_letter<W>(c.klo, c.khi, (uint32_t)i, true, c.k, (int)j);
  const uint32_t* ord = c.order + c.ubase[i];
  const uint32_t o = j < (uint64_t)c.k ? ord[0] : ord[j - c.k + 1];
  const int q = j < (uint64_t)c.k ? (int)j : c.k - 1;
  return oriented_letter<W>(c.klo, c.khi, o & 0x7fffffffu, !(o & 0x80000000u), c.k, q);
}

// A wave per vertex; lanes over the letters.
template <int W>
__global__ void __launch_bounds__(kBlock) k_fasta_write(VertexCtx c, const uint64_t* __restrict__ off,
                                                       uint8_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t i = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); i < c.n; i += nw) {
    uint8_t* p = out + off[i];
    const uint32_t key = c.merged ? c.urep_node[i] : (uint32_t)i;
    const uint64_t len = fasta_seq_len<W>(c, i);
    if (lane < 6) p[lane] = ">node_"[lane];
    for (int j = lane; j < c.k; j += 64) p[6 + j] = kAcgt[oriented_letter<W>(c.klo, c.khi, key, true, c.k, j)];
    if (lane == 0) p[6 + c.k] = '\n';
    uint8_t* q = p + 6 + c.k + 1;
    for (uint64_t j = lane; j < len; j += 64) q[j] = kAcgt[fasta_letter<W>(c, i, j)];
    if (lane == 0) q[len] = '\n';
  }
}

template <int W>
static VertexCtx vertex_ctx(gx_build* b, int merged) {
  VertexCtx c{};
  c.merged = merged;
  c.k = b->cfg.k;
  c.klo = b->klo.as<uint64_t>();
  c.khi = W == 2 ? b->khi.as<uint64_t>() : nullptr;
  if (!merged) {
    c.n = b->nnodes;
    c.cnt = b->cnt.as<uint32_t>();
    c.msk = b->msk.as<uint32_t>();
    c.hcu = b->hcu.as<uint32_t>();
    c.hcf = b->hcf.as<uint32_t>();
  } else {
    c.n = b->u_count;
    c.ulen = b->u_len.as<uint64_t>();
    c.ucov = b->u_cov.as<unsigned long long>();
    c.uedge = b->u_edge.as<uint32_t>();
    c.uecnt = b->u_ecnt.as<uint8_t>();
    c.urep_node = b->u_repnode.as<uint32_t>();
    c.hu = b->u_hu.as<uint32_t>();
    c.hf = b->u_hf.as<uint32_t>();
    c.ubase = b->u_base.as<uint64_t>();
    c.order = b->u_order.as<uint32_t>();
  }
  return c;
}

template <int W>
static void graph_stats(gx_build* b, int merged) {
  hipStream_t st = b->st;
  const unsigned G = b->grid;
  const VertexCtx c = vertex_ctx<W>(b, merged);
  const uint64_t n = c.n;
  unsigned long long* acc = buf<unsigned long long>(b->s_acc, 4 + NTOT);
  GX_HIP(hipMemsetAsync(acc, 0, (4 + NTOT) * 8, st));
  if (n) {
    hipLaunchKernelGGL(k_stat_totals<W>, dim3(grid_for(n, kBlock, G)), dim3(kBlock), 0, st, c, acc + 4);
    GX_HIP(hipGetLastError());
  }
  b->st_tot.assign(NTOT, 0);
  GX_HIP(hipMemcpyAsync(b->st_tot.data(), acc + 4, NTOT * 8, hipMemcpyDeviceToHost, st));
  uint64_t* k1 = buf<uint64_t>(b->s_k1, n + 1);
  uint64_t* k2 = buf<uint64_t>(b->s_k2, n + 1);
  uint32_t* v1 = buf<uint32_t>(b->s_v1, n + 1);
  uint32_t* v2 = buf<uint32_t>(b->s_v2, n + 1);
  uint32_t* flag = buf<uint32_t>(b->s_flag, n + 1);
  uint64_t* pos = buf<uint64_t>(b->s_pos, n + 1);
  b->st_bins.assign(NSTAT, {});
  b->st_sum.assign(NSTAT, 0);
  b->st_max.assign(NSTAT, 0);
  for (int s = 0; s < NSTAT; s++) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, st, reinterpret_cast<uint64_t*>(acc) + 3, ~0ull);
    GX_HIP(hipMemsetAsync(acc, 0, 24, st));
    if (n) {
      hipLaunchKernelGGL(k_stat_values<W>, dim3(grid_for(n, kBlock, G)), dim3(kBlock), 0, st, c, s, k1, flag, acc);
      GX_HIP(hipGetLastError());
    }
    unsigned long long a[4];
    GX_HIP(hipMemcpyAsync(a, acc, 32, hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    const uint64_t m = a[0];
    b->st_sum[s] = (int64_t)a[1];
    const int64_t mx = m ? (int64_t)(a[2] - 0x8000000000000000ull) : 0;
    b->st_max[s] = mx > 0 ? mx : 0;  // the reference's maximum counter starts at 0
    if (!m) continue;
    const uint64_t lo = a[3], range = a[2] - a[3];
    int nbytes = 0;
    while (nbytes < 8 && (range >> (8 * nbytes)) != 0) nbytes++;
    exclusive_scan<uint32_t, uint64_t>(flag, pos, n, scratch_for(b, scan_scratch_bytes<uint64_t>(n)), st);
    hipLaunchKernelGGL(k_stat_compact, dim3(grid_for(n, kBlock, G)), dim3(kBlock), 0, st, n, k1, flag, pos, lo, k2,
                       v2);
    GX_HIP(hipGetLastError());
    bool f = nbytes ? radix_sort<uint64_t>(k2, v2, k1, v1, m, 0, nbytes,
                                           scratch_for(b, radix_scratch_bytes<uint64_t>(m)), st)
                    : false;
    uint64_t* ks = f ? k1 : k2;
    hipLaunchKernelGGL(k_bin_starts, dim3(grid_for(m, kBlock, G)), dim3(kBlock), 0, st, m, ks, flag);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint32_t, uint64_t>(flag, pos, m, scratch_for(b, scan_scratch_bytes<uint64_t>(m)), st);
    const uint64_t nbin = read1(st, pos + m);
    int64_t* bval = buf<int64_t>(b->s_bval, nbin + 1);
    uint64_t* bst = buf<uint64_t>(b->s_bst, nbin + 1);
    hipLaunchKernelGGL(k_bin_write, dim3(grid_for(m, kBlock, G)), dim3(kBlock), 0, st, m, ks, lo, flag, pos, bval,
                       bst);
    GX_HIP(hipGetLastError());
    std::vector<int64_t> hv(nbin);
    std::vector<uint64_t> hs(nbin);
    GX_HIP(hipMemcpyAsync(hv.data(), bval, nbin * 8, hipMemcpyDeviceToHost, st));
    GX_HIP(hipMemcpyAsync(hs.data(), bst, nbin * 8, hipMemcpyDeviceToHost, st));
    GX_HIP(hipStreamSynchronize(st));
    auto& bins = b->st_bins[s];
    bins.resize(nbin);
    for (uint64_t q = 0; q < nbin; q++) bins[q] = {hv[q], (q + 1 < nbin ? hs[q + 1] : m) - hs[q]};
  }
  b->st_graph = merged;
  b->reclaim.dead.insert(b->reclaim.dead.end(),
                         {&b->s_k1, &b->s_k2, &b->s_v1, &b->s_v2, &b->s_flag, &b->s_pos, &b->s_bval, &b->s_bst});
}

template <int W>
static uint64_t fasta_bytes(gx_build* b, int merged, uint8_t* dst, size_t cap) {
  hipStream_t st = b->st;
  const unsigned G = b->grid;
  const VertexCtx c = vertex_ctx<W>(b, merged);
  const uint64_t n = c.n;
  uint64_t* sz = buf<uint64_t>(b->s_k1, n + 1);
  uint64_t* off = buf<uint64_t>(b->s_k2, n + 1);
  if (n) {
    hipLaunchKernelGGL(k_fasta_size<W>, dim3(grid_for(n, kBlock, G)), dim3(kBlock), 0, st, c, sz);
    GX_HIP(hipGetLastError());
  }
  exclusive_scan<uint64_t, uint64_t>(sz, off, n, scratch_for(b, scan_scratch_bytes<uint64_t>(n)), st);
  const uint64_t total = read1(st, off + n);
  if (!dst) return total;
  if (cap < total) throw GxError(GX_EINVAL, "destination too small");
  uint8_t* out = buf<uint8_t>(b->s_out, total + 1);
  if (n) {
    hipLaunchKernelGGL(k_fasta_write<W>, dim3(grid_for(n, kBlock / 64, 1u << 16)), dim3(kBlock), 0, st, c, off, out);
    GX_HIP(hipGetLastError());
  }
  if (total) GX_HIP(hipMemcpyAsync(dst, out, total, hipMemcpyDeviceToHost, st));
  GX_HIP(hipStreamSynchronize(st));
  return total;
}

extern "C" {

// Checks of a configuration (GenomixJobConf.java:417-423 for k); returns a
// GX_* code and sets msg.
static int check_cfg(const gx_build_cfg* cfg, std::string& msg) {
  if (cfg->k < 3) { msg = "kmerLength must be at least 3!"; return GX_EINVAL; }
  if (cfg->k > 63) { msg = "k > 63 is not supported"; return GX_EINVAL; }
  if (cfg->max_readids_per_edge < 0) { msg = "maxReadIDsPerEdge must be non-negative!"; return GX_EINVAL; }
  if (cfg->nparts < 1) { msg = "nparts must be >= 1"; return GX_EINVAL; }
  if (cfg->nparts > 65536) { msg = "nparts above 65536 is not supported"; return GX_ENOTSUP; }
  if (cfg->edge_order != GX_EDGE_ORDER_CANONICAL && cfg->edge_order != GX_EDGE_ORDER_JDK7) {
    msg = "unknown edge_order";
    return GX_EINVAL;
  }
  return GX_OK;
}

int gx_build_create(const gx_build_cfg* cfg, gx_build** out) {
  g_create_err.clear();
  if (!cfg || !out) { g_create_err = "null argument"; return GX_EINVAL; }
  if (int rc = check_cfg(cfg, g_create_err)) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    g_create_err = "no HIP device available: the graph build runs on the GPU only";
    return GX_EDEVICE;
  }
  if (cfg->device < 0 || cfg->device >= ndev) { g_create_err = "bad device ordinal"; return GX_EINVAL; }
  gx_build* b = new (std::nothrow) gx_build;
  if (!b) return GX_ENOMEM;
  b->cfg = *cfg;
  b->W = cfg->k <= 31 ? 1 : 2;
  b->nb = (cfg->k + 3) / 4;
  int rc = guarded(b, [&] {
    GX_HIP(hipSetDevice(cfg->device));
    GX_HIP(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
  });
  if (rc) {
    g_create_err = b->err;
    delete b;
    return rc;
  }
  *out = b;
  return GX_OK;
}

int gx_build_add_reads(gx_build* b, const char* buf_, size_t len, int fmt) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (len && !buf_) throw GxError(GX_EINVAL, "null buffer");
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    append_text(b, buf_, len, hipMemcpyHostToDevice, fmt);
  });
}

int gx_build_add_reads_device(gx_build* b, const void* dev_buf, size_t len, int fmt) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (len && !dev_buf) throw GxError(GX_EINVAL, "null buffer");
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    append_text(b, dev_buf, len, hipMemcpyDeviceToDevice, fmt);
  });
}

int gx_build_add_fastq_pair(gx_build* b, const char* mate1, size_t n1, const char* mate2, size_t n2) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if ((n1 && !mate1) || (n2 && !mate2)) throw GxError(GX_EINVAL, "null buffer");
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    set_format(b, kFmtFastqPair);
    append_to(b, b->text, b->text_len, mate1, n1, hipMemcpyHostToDevice);
    append_to(b, b->text2, b->text2_len, mate2, n2, hipMemcpyHostToDevice);
  });
}

int gx_build_readids(gx_build* b, char* dst, size_t cap, size_t* written) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->finished) throw GxError(GX_ESTATE, "readids before finish");
    if (b->fmt != GX_FMT_FASTQ && b->fmt != kFmtFastqPair)
      throw GxError(GX_ESTATE, "the handle was given readids text, not FASTQ");
    if (!dst) {
      if (!written) throw GxError(GX_EINVAL, "null argument");
      *written = b->rtext_len;
      return;
    }
    if (cap < b->rtext_len) throw GxError(GX_EINVAL, "destination too small");
    GX_HIP(hipSetDevice(b->cfg.device));
    if (b->rtext_len) GX_HIP(hipMemcpyAsync(dst, b->rtext.p, b->rtext_len, hipMemcpyDeviceToHost, b->st));
    GX_HIP(hipStreamSynchronize(b->st));
    if (written) *written = b->rtext_len;
  });
}

int gx_build_finish(gx_build* b) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (b->finished) throw GxError(GX_ESTATE, "finish called twice; call gx_build_reset first");
    GX_HIP(hipSetDevice(b->cfg.device));
    if (b->text.cap == 0) b->text.ensure(64);
    if (b->fmt == GX_FMT_FASTQ || b->fmt == kFmtFastqPair) {
      fastq_convert(b);
      b->src = b->rtext.as<uint8_t>();
      b->src_len = b->rtext_len;
    } else {
      b->src = b->text.as<uint8_t>();
      b->src_len = b->text_len;
    }
    b->unitigs_done = false;
    b->st_graph = -1;
    results_dead(b);
    ReclaimScope rs(&b->reclaim, b->st);
    b->reclaim.allocs = b->reclaim.freed = 0;
    if (b->W == 1) run_build<1>(b);
    else run_build<2>(b);
    b->stats[14] = b->reclaim.allocs;
    b->stats[15] = b->reclaim.freed;
    b->finished = true;
  });
}

int gx_build_part_info(gx_build* b, int part, uint64_t* nrec, uint64_t* nbytes) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->finished) throw GxError(GX_ESTATE, "part_info before finish");
    if (part < 0 || part >= b->cfg.nparts) throw GxError(GX_EINVAL, "no such partition");
    if (nrec) *nrec = b->part_rec[part + 1] - b->part_rec[part];
    if (nbytes) *nbytes = b->part_off[part + 1] - b->part_off[part];
  });
}

int gx_build_emit_part(gx_build* b, int part, uint8_t* dst, size_t cap, size_t* written) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->finished) throw GxError(GX_ESTATE, "emit before finish");
    if (b->streamed) throw GxError(GX_ESTATE, "the records went to the sink");
    if (part < 0 || part >= b->cfg.nparts) throw GxError(GX_EINVAL, "no such partition");
    uint64_t off = b->part_off[part], n = b->part_off[part + 1] - off;
    if (cap < n) throw GxError(GX_EINVAL, "destination too small");
    if (n && !dst) throw GxError(GX_EINVAL, "null destination");
    GX_HIP(hipSetDevice(b->cfg.device));
    if (n) GX_HIP(hipMemcpyAsync(dst, b->out.as<uint8_t>() + off, n, hipMemcpyDeviceToHost, b->st));
    GX_HIP(hipStreamSynchronize(b->st));
    if (written) *written = n;
  });
}

int gx_build_part_device(gx_build* b, int part, const void** dev_ptr, uint64_t* nbytes) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->finished) throw GxError(GX_ESTATE, "part_device before finish");
    if (b->streamed) throw GxError(GX_ESTATE, "the records went to the sink");
    if (part < 0 || part >= b->cfg.nparts) throw GxError(GX_EINVAL, "no such partition");
    if (!dev_ptr || !nbytes) throw GxError(GX_EINVAL, "null argument");
    *dev_ptr = b->out.as<uint8_t>() + b->part_off[part];
    *nbytes = b->part_off[part + 1] - b->part_off[part];
  });
}

int gx_build_stats(gx_build* b, uint64_t* out, int n) {
  if (!b || !out) return GX_EINVAL;
  b->stats[10] = b->u_rounds;
  b->stats[12] = b->u_succ_reads;
  b->stats[13] = b->u_walkers;
  for (int i = 0; i < n && i < 16; i++) out[i] = b->stats[i];
  return GX_OK;
}

int gx_clean_bad_coverage(gx_build* b, float min_cov, float max_cov) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->unitigs_done) throw GxError(GX_ESTATE, "clean_bad_coverage before unitig_run");
    if (b->nnodes >= kMaxUnitigNodes) throw GxError(GX_ENOTSUP, "graph too large for the unitig pass");
    if (!(min_cov >= 0.f) || !(max_cov >= 0.f))  // GenomixJobConf.java:442-447
      throw GxError(GX_EINVAL, "coverage bounds must not be negative");
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    b->unitigs_done = false;
    b->st_graph = -1;
    const CleanArgs ca{CLEAN_COVERAGE, min_cov, max_cov, 0, b->cfg.k};
    if (b->W == 1) clean_graph<1>(b, ca);
    else clean_graph<2>(b, ca);
    b->unitigs_done = true;
  });
}

int gx_clean_tips(gx_build* b, int max_len) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (max_len >= 0 && max_len < b->cfg.k)  // GenomixJobConf.java:449-450
      throw GxError(GX_EINVAL, "tipRemove_maxLength must be at least as long as kmerLength");
    if (!b->unitigs_done) throw GxError(GX_ESTATE, "clean_tips before unitig_run");
    if (b->nnodes >= kMaxUnitigNodes) throw GxError(GX_ENOTSUP, "graph too large for the unitig pass");
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    b->unitigs_done = false;
    b->st_graph = -1;
    const CleanArgs ca{CLEAN_TIP, 0.f, 0.f, (uint64_t)(max_len < 0 ? b->cfg.k : max_len), b->cfg.k};
    if (b->W == 1) clean_graph<1>(b, ca);
    else clean_graph<2>(b, ca);
    b->unitigs_done = true;
  });
}

int gx_clean_bubbles(gx_build* b, float max_dissimilarity, int max_len) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    // GenomixJobConf.java:429-432 (NaN passes, as there, and merges nothing)
    if (max_dissimilarity < 0.f || max_dissimilarity > 1.f)
      throw GxError(GX_EINVAL, "bubbleMerge_maxDissimilarity must lie in [0, 1]");
    const uint64_t maxlen = max_len < 0 ? 5ull * (uint64_t)b->cfg.k : (uint64_t)max_len;  // :481-482
    if (maxlen > kBubMaxLetters + 1)
      throw GxError(GX_ENOTSUP, "bubbles longer than " + std::to_string(kBubMaxLetters) + " letters");
    if (!b->unitigs_done) throw GxError(GX_ESTATE, "clean_bubbles before unitig_run");
    if (b->nnodes >= kMaxUnitigNodes) throw GxError(GX_ENOTSUP, "graph too large for the unitig pass");
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    b->unitigs_done = false;
    b->st_graph = -1;
    if (b->W == 1) bubble_graph<1>(b, max_dissimilarity, maxlen);
    else bubble_graph<2>(b, max_dissimilarity, maxlen);
    b->unitigs_done = true;
  });
}

// Densities of commons-math3 3.2 (the version genomix-driver/pom.xml:367-368
// pins): ExponentialDistribution.density and NormalDistribution.density.
static double mix_exp_density(double mean, double x) { return x < 0 ? 0.0 : std::exp(-x / mean) / mean; }
static double mix_norm_density(double mu, double sd, double x) {
  const double z = (x - mu) / sd;
  return std::exp(-0.5 * z * z) / (sd * 2.5066282746310002);  // sd * sqrt(2 pi)
}

// FittingMixture.fittingMixture (FittingMixture.java:89-219) over binned data:
// value x[i] stands for w[i] equal data points, so every sum over the data is
// a w-weighted sum over the bins.  EM of an exponential plus a normal
// component from mean 5, normal 20 / 5 and priors 1/2; an E step, then
// `iters` rounds of M step (means, sd floored at 1e-9f) and E step, each E
// step ending with the priors' update.  out: cutoff, exponential mean,
// normal mean, normal sd, exponential prior, normal prior.
static void fit_mixture(const double* x, const uint64_t* w, size_t n, double max, int iters, double* out) {
  double total = 0;
  for (size_t i = 0; i < n; i++) total += (double)w[i];
  if (n == 0 || total == 0 || max == 0) throw GxError(GX_EINVAL, "No information for coverage!");
  if (iters < 0) throw GxError(GX_EINVAL, "negative iteration count");
  std::vector<double> me(n), mn(n);
  double emean = 5, nmean = 20, nsd = 5, pe = 0.5, pn = 0.5, se = 0, sn = 0;
  auto estep = [&] {
    se = sn = 0;
    for (size_t i = 0; i < n; i++) {
      double a = mix_exp_density(emean, x[i]) * pe;
      if (a == 0) a = 0.000000001;
      double c = mix_norm_density(nmean, nsd, x[i]) * pn;
      if (c == 0) c = 0.000000001;
      const double s = a + c;
      me[i] = a / s;
      mn[i] = c / s;
      se += (double)w[i] * me[i];
      sn += (double)w[i] * mn[i];
    }
    pe = se / (se + sn);
    pn = sn / (se + sn);
  };
  estep();
  for (int it = 0; it < iters; it++) {
    double a = 0, c = 0;
    for (size_t i = 0; i < n; i++) {
      a += (double)w[i] * me[i] * x[i];
      c += (double)w[i] * mn[i] * x[i];
    }
    emean = a / se;
    // ExponentialDistribution's constructor rejects a mean <= 0
    // (NotStrictlyPositiveException).
    if (emean <= 0) throw GxError(GX_EINVAL, "exponential mean must be positive");
    nmean = c / sn;
    double v = 0;
    for (size_t i = 0; i < n; i++) v += (double)w[i] * mn[i] * (x[i] - nmean) * (x[i] - nmean);
    nsd = std::sqrt(v / sn);
    if (nsd == 0) nsd = (double)0.000000001f;
    estep();
  }
  double cut = 0;
  for (double cov = 1; cov < max; cov++)
    if (pe * mix_exp_density(emean, cov) < pn * mix_norm_density(nmean, nsd, cov)) {
      cut = cov;
      break;
    }
  out[0] = cut;
  out[1] = emean;
  out[2] = nmean;
  out[3] = nsd;
  out[4] = pe;
  out[5] = pn;
}

static void check_graph(gx_build* b, int graph) {
  if (graph != GX_GRAPH_NODES && graph != GX_GRAPH_UNITIGS) throw GxError(GX_EINVAL, "unknown graph");
  if (!b->finished) throw GxError(GX_ESTATE, "statistics before finish");
  if (!b->graph_held)
    throw GxError(GX_ENOTSUP, "the node graph of this key-range build is not held in HBM (its records went to a sink, "
                              "or it has 2^32 or more nodes)");
  if (graph == GX_GRAPH_UNITIGS && !b->unitigs_done) throw GxError(GX_ESTATE, "no unitig pass to report on");
}

int gx_graph_stats(gx_build* b, int graph) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    check_graph(b, graph);
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    if (b->W == 1) graph_stats<1>(b, graph);
    else graph_stats<2>(b, graph);
  });
}

int gx_graph_stat_totals(gx_build* b, uint64_t* out, int n) {
  if (!b || !out) return GX_EINVAL;
  return guarded(b, [&] {
    if (b->st_graph < 0) throw GxError(GX_ESTATE, "gx_graph_stats first");
    for (int i = 0; i < n && i < (int)b->st_tot.size(); i++) out[i] = b->st_tot[i];
  });
}

int gx_graph_stat(gx_build* b, int stat, int64_t* total, int64_t* maximum, uint64_t* nbins) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (b->st_graph < 0) throw GxError(GX_ESTATE, "gx_graph_stats first");
    if (stat < 0 || stat >= NSTAT) throw GxError(GX_EINVAL, "unknown statistic");
    if (total) *total = b->st_sum[stat];
    if (maximum) *maximum = b->st_max[stat];
    if (nbins) *nbins = b->st_bins[stat].size();
  });
}

int gx_graph_stat_bins(gx_build* b, int stat, int64_t* values, uint64_t* counts, size_t cap) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (b->st_graph < 0) throw GxError(GX_ESTATE, "gx_graph_stats first");
    if (stat < 0 || stat >= NSTAT) throw GxError(GX_EINVAL, "unknown statistic");
    const auto& bins = b->st_bins[stat];
    if (cap < bins.size()) throw GxError(GX_EINVAL, "destination too small");
    for (size_t q = 0; q < bins.size(); q++) {
      if (values) values[q] = bins[q].first;
      if (counts) counts[q] = bins[q].second;
    }
  });
}

int gx_coverage_cutoff(gx_build* b, int iterations, double* out) {
  if (!b || !out) return GX_EINVAL;
  return guarded(b, [&] {
    if (b->st_graph < 0) throw GxError(GX_ESTATE, "gx_graph_stats first");
    const auto& bins = b->st_bins[ST_COV];
    // GraphStatistics.getCoverageStats / getMaxCoverage
    // (GraphStatistics.java:165-199): the counters' values and the largest.
    std::vector<double> x(bins.size());
    std::vector<uint64_t> w(bins.size());
    double mx = 0;
    for (size_t q = 0; q < bins.size(); q++) {
      x[q] = (double)bins[q].first;
      w[q] = bins[q].second;
      mx = std::max(mx, x[q]);
    }
    if (bins.empty() || mx == 0) throw GxError(GX_ESTATE, "No information for coverage!");
    fit_mixture(x.data(), w.data(), x.size(), mx, iterations, out);
  });
}

int gx_fit_mixture(const double* values, const uint64_t* counts, size_t n, double max, int iterations, double* out) {
  g_create_err.clear();
  if ((n && (!values || !counts)) || !out) {
    g_create_err = "null argument";
    return GX_EINVAL;
  }
  try {
    fit_mixture(values, counts, n, max, iterations, out);
    return GX_OK;
  } catch (const GxError& e) {
    g_create_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_create_err = "host allocation failed";
    return GX_ENOMEM;
  }
}

int gx_dump_fasta(gx_build* b, int graph, uint8_t* dst, size_t cap, size_t* written) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    check_graph(b, graph);
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    const uint64_t n = b->W == 1 ? fasta_bytes<1>(b, graph, dst, cap) : fasta_bytes<2>(b, graph, dst, cap);
    if (written) *written = n;
  });
}

int gx_unitig_run(gx_build* b) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->finished) throw GxError(GX_ESTATE, "unitig_run before finish");
    if (!b->graph_held)
      throw GxError(GX_ENOTSUP, "the unitig pass needs the whole graph in HBM; this build of " +
                                    std::to_string(b->nranges) +
                                    " key ranges handed them to a record sink or holds 2^32 or more nodes");
    GX_HIP(hipSetDevice(b->cfg.device));
    ReclaimScope rs(&b->reclaim, b->st);
    b->st_graph = -1;
    if (b->W == 1) run_unitigs<1>(b);
    else run_unitigs<2>(b);
    b->unitigs_done = true;
  });
}

int gx_unitig_set_ranking(gx_build* b, int mode) {
  if (!b) return GX_EINVAL;
  if (mode != GX_RANK_AUTO && mode != GX_RANK_POINTER_JUMPING) {
    b->err = "unknown ranking mode";
    return GX_EINVAL;
  }
  b->no_walk = mode == GX_RANK_POINTER_JUMPING;
  return GX_OK;
}

int gx_unitig_limits(uint64_t* max_nodes, uint64_t* max_head_offset) {
  if (max_nodes) *max_nodes = kMaxUnitigNodes;
  if (max_head_offset) *max_head_offset = kMaxHeadOffset;
  return GX_OK;
}

int gx_unitig_info(gx_build* b, uint64_t* nrec, uint64_t* nbytes) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->unitigs_done) throw GxError(GX_ESTATE, "unitig_info before unitig_run");
    if (nrec) *nrec = b->u_count;
    if (nbytes) *nbytes = b->u_bytes;
  });
}

int gx_unitig_emit(gx_build* b, uint8_t* dst, size_t cap, size_t* written) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->unitigs_done) throw GxError(GX_ESTATE, "unitig_emit before unitig_run");
    if (cap < b->u_bytes) throw GxError(GX_EINVAL, "destination too small");
    GX_HIP(hipSetDevice(b->cfg.device));
    if (b->u_bytes) GX_HIP(hipMemcpyAsync(dst, b->u_out.p, b->u_bytes, hipMemcpyDeviceToHost, b->st));
    GX_HIP(hipStreamSynchronize(b->st));
    if (written) *written = b->u_bytes;
  });
}

int gx_build_set_tuning(gx_build* b, int grid_blocks, float table_slack) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (grid_blocks < 64 || grid_blocks > 65536) throw GxError(GX_EINVAL, "grid_blocks must be in [64, 65536]");
    if (!(table_slack >= 1.05f && table_slack <= 16.f)) throw GxError(GX_EINVAL, "table_slack must be in [1.05, 16]");
    b->grid = (unsigned)grid_blocks;
    b->slack = table_slack;
  });
}

int gx_build_set_method(gx_build* b, int method) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (method != GX_AGG_SORT && method != GX_AGG_HASH) throw GxError(GX_EINVAL, "unknown aggregation method");
    b->method = method;
  });
}

int gx_build_set_stream(gx_build* b, uint64_t range_occurrences, gx_part_sink sink, void* user) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    b->range_occ = range_occurrences;
    b->sink = sink;
    b->sink_user = user;
  });
}

int gx_build_set_config(gx_build* b, const gx_build_cfg* cfg) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!cfg) throw GxError(GX_EINVAL, "null argument");
    std::string msg;
    if (int rc = check_cfg(cfg, msg)) throw GxError(rc, msg);
    if (cfg->device != b->cfg.device) throw GxError(GX_EINVAL, "a handle stays on its device");
    b->cfg = *cfg;
    b->W = cfg->k <= 31 ? 1 : 2;
    b->nb = (cfg->k + 3) / 4;
    b->text_len = b->text2_len = b->rtext_len = 0;
    b->fmt = -1;
    b->finished = false;
    b->unitigs_done = false;
    b->out_bytes = 0;
    b->nnodes = 0;
    b->st_graph = -1;
  });
}

int gx_build_set_sizer(gx_build* b, gx_mlp* m) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (m) {
      int din = 0, dout = 0;
      gx_mlp_dims(m, &din, &dout);
      if (din != GX_SIZER_FEATURES || dout != GX_SIZER_OUTPUTS)
        throw GxError(GX_EINVAL, "sizing network must map 72 features to 2 outputs");
      if (gx_mlp_device_of(m) != b->cfg.device) throw GxError(GX_EINVAL, "sizing network lives on another device");
    }
    b->sizer = m;
  });
}

int gx_build_features(gx_build* b, float* feat, int n, uint8_t* regs, double* prediction) {
  if (!b) return GX_EINVAL;
  return guarded(b, [&] {
    if (!b->finished) throw GxError(GX_ESTATE, "features before finish");
    if (feat) {
      if (n < GX_SIZER_FEATURES) throw GxError(GX_EINVAL, "feature buffer too small");
      memcpy(feat, b->feat, sizeof(b->feat));
    }
    if (regs) memcpy(regs, b->regs_h, sizeof(b->regs_h));
    if (prediction) {
      prediction[0] = b->pred[0];
      prediction[1] = b->pred[1];
    }
  });
}

int gx_build_set_profiling(gx_build* b, int on) {
  if (!b) return GX_EINVAL;
  b->prof = on != 0;
  return GX_OK;
}

int gx_build_phase_times(gx_build* b, double* ms, int n) {
  if (!b || !ms) return GX_EINVAL;
  for (int i = 0; i < n && i < NPH; i++) ms[i] = b->phase_ms[i];
  return NPH;
}

const char* gx_build_phase_name(int i) { return (i >= 0 && i < NPH) ? kPhaseNames[i] : ""; }

int gx_build_reset(gx_build* b) {
  if (!b) return GX_EINVAL;
  b->text_len = 0;
  b->text2_len = 0;
  b->rtext_len = 0;
  b->fmt = -1;
  b->finished = false;
  b->unitigs_done = false;
  b->out_bytes = 0;
  b->nnodes = 0;
  b->st_graph = -1;  // statistics describe the graph that is gone
  return GX_OK;
}

const char* gx_last_error(const gx_build* b) { return b ? b->err.c_str() : g_create_err.c_str(); }

void gx_build_destroy(gx_build* b) {
  if (!b) return;
  (void)hipSetDevice(b->cfg.device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->st2) (void)hipStreamSynchronize(b->st2);
  if (b->st) (void)hipStreamDestroy(b->st);
  if (b->st2) (void)hipStreamDestroy(b->st2);
  for (hipEvent_t e : {b->ev_fork, b->ev_join})
    if (e) (void)hipEventDestroy(e);
  for (auto& e : b->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : b->evpj)
    if (e) (void)hipEventDestroy(e);
  delete b;
}

}  // extern "C"
