// k-mer arithmetic on the device.
//
// A k-mer s_0..s_{k-1} is the integer V = sum code(s_i) * 4^i (letter 0 in
// the low bits).  Its big-endian bytes over nb = ceil(k/4) bytes are exactly
// the Java Kmer/VKmer packing (Kmer.java:225-242), so numeric order of V is
// the byte order BinaryComparable.compareTo uses for the canonical choice
// (ReadsKeyValueParserFactory.java:163,181).  Keys up to k = 31 live in one
// 64-bit word (W = 1); up to k = 63 in two (W = 2, lo = bits 0..63).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gx {

struct KV {
  uint64_t lo, hi;
};

__device__ __host__ inline uint64_t rev2_64(uint64_t x) {
  x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
  return __builtin_bswap64(x);
}

// Mask of the 2k valid bits.
template <int W>
__device__ __host__ inline KV kmask(int k) {
  KV m;
  int bits = 2 * k;
  if (W == 1) {
    m.lo = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    m.hi = 0;
  } else {
    if (bits >= 128) { m.lo = ~0ull; m.hi = ~0ull; }
    else if (bits > 64) { m.lo = ~0ull; m.hi = (1ull << (bits - 64)) - 1; }
    else { m.lo = bits == 64 ? ~0ull : ((1ull << bits) - 1); m.hi = 0; }
  }
  return m;
}

// Reverse complement: complement every 2-bit code, reverse the order of the
// k codes (Kmer.setReversedFromStringBytes, Kmer.java:253-272).
template <int W>
__device__ __host__ inline KV revcomp(KV v, int k) {
  KV r;
  if (W == 1) {
    r.lo = rev2_64(~v.lo) >> (64 - 2 * k);
    r.hi = 0;
  } else {
    uint64_t a = rev2_64(~v.hi);  // becomes the low word of the 128-bit reversal
    uint64_t b = rev2_64(~v.lo);  // becomes the high word
    int sh = 128 - 2 * k;        // 2 <= sh <= 64 for 32 <= k <= 63
    if (sh >= 64) {
      r.lo = b >> (sh - 64);
      r.hi = 0;
    } else {
      r.lo = (a >> sh) | (b << (64 - sh));
      r.hi = b >> sh;
    }
  }
  return r;
}

template <int W>
__device__ __host__ inline bool kle(KV a, KV b) {
  if (W == 1) return a.lo <= b.lo;
  return a.hi < b.hi || (a.hi == b.hi && a.lo <= b.lo);
}
template <int W>
__device__ __host__ inline bool keq(KV a, KV b) {
  if (W == 1) return a.lo == b.lo;
  return a.hi == b.hi && a.lo == b.lo;
}

// Successor X[1:]+c and predecessor c+X[:-1] of a k-mer.
template <int W>
__device__ __host__ inline KV succ_of(KV v, int k, uint32_t c) {
  KV r;
  if (W == 1) {
    r.lo = (v.lo >> 2) | ((uint64_t)c << (2 * k - 2));
    r.hi = 0;
  } else {
    r.lo = (v.lo >> 2) | (v.hi << 62);
    r.hi = v.hi >> 2;
    int p = 2 * k - 2;
    if (p >= 64) r.hi |= (uint64_t)c << (p - 64);
    else r.lo |= (uint64_t)c << p;
  }
  return r;
}
template <int W>
__device__ __host__ inline KV pred_of(KV v, int k, uint32_t c) {
  KV r, m = kmask<W>(k);
  if (W == 1) {
    r.lo = ((v.lo << 2) | c) & m.lo;
    r.hi = 0;
  } else {
    r.hi = ((v.hi << 2) | (v.lo >> 62)) & m.hi;
    r.lo = ((v.lo << 2) | c) & m.lo;
  }
  return r;
}

template <int W>
__device__ __host__ inline KV canon(KV v, int k, bool* fwd) {
  KV r = revcomp<W>(v, k);
  bool f = kle<W>(v, r);
  if (fwd) *fwd = f;
  return f ? v : r;
}

__device__ __host__ inline uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

template <int W>
__device__ __host__ inline uint64_t khash(KV v) {
  return W == 1 ? mix64(v.lo) : mix64(v.lo ^ mix64(v.hi + 0x9e3779b97f4a7c15ull));
}

// Byte j (0 = most significant) of the nb-byte big-endian form of V.
template <int W>
__device__ __host__ inline uint8_t be_byte(KV v, int nb, int j) {
  int q = nb - 1 - j;  // byte index from the least significant end
  return (uint8_t)(q < 8 ? (v.lo >> (8 * q)) : (v.hi >> (8 * (q - 8))));
}

// Java int hash of the nb key bytes (KmerPartitionComputerFactory.hashBytes,
// KmerPartitionComputerFactory.java:31-36).
template <int W>
__device__ __host__ inline int32_t java_hash_key(KV v, int nb) {
  uint32_t h = 1;
  for (int j = 0; j < nb; j++) h = h * 31u + (uint32_t)(int32_t)(int8_t)be_byte<W>(v, nb, j);
  return (int32_t)h;
}

// VKmer.hashCode (VKmer.java:394-396): hash over BE32(k) || key bytes.
template <int W>
__device__ __host__ inline int32_t java_hash_vkmer(KV v, int k, int nb) {
  uint32_t h = 1;
  h = h * 31u + (uint32_t)(int32_t)(int8_t)((k >> 24) & 0xff);
  h = h * 31u + (uint32_t)(int32_t)(int8_t)((k >> 16) & 0xff);
  h = h * 31u + (uint32_t)(int32_t)(int8_t)((k >> 8) & 0xff);
  h = h * 31u + (uint32_t)(int32_t)(int8_t)(k & 0xff);
  for (int j = 0; j < nb; j++) h = h * 31u + (uint32_t)(int32_t)(int8_t)be_byte<W>(v, nb, j);
  return (int32_t)h;
}

// JDK 7 HashMap.hash(int) supplemental spreading.
__device__ __host__ inline int32_t jdk7_spread(int32_t h) {
  uint32_t x = (uint32_t)h;
  x ^= (x >> 20) ^ (x >> 12);
  return (int32_t)(x ^ (x >> 7) ^ (x >> 4));
}

}  // namespace gx
