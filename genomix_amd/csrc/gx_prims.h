// Device-wide primitives for the graph build, written for gfx950 (wave64):
// an exclusive prefix sum and a stable LSD radix sort.  Both are plain
// integer, HBM-bound kernels; tiles are sized for 256-thread workgroups.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gx {

#define GX_HIP(call)                                                                               \
  do {                                                                                             \
    hipError_t _e = (call);                                                                        \
    if (_e != hipSuccess)                                                                          \
      throw ::gx::HipError(std::string(#call) + ": " + hipGetErrorString(_e), (int)_e);            \
  } while (0)

struct HipError : std::runtime_error {
  int code;
  HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

struct DevBuf;

// Buffers that the running build reads no more.  They stay allocated for the
// next build on the handle and are freed only when an allocation would
// otherwise fail; ensure() of a listed buffer takes it off the list.
struct ReclaimList {
  hipStream_t st = nullptr;
  hipStream_t st2 = nullptr;  // a second stream of the same handle (drained too before a free)
  std::vector<DevBuf*> dead;
  uint64_t allocs = 0;  // device allocations made while the list was installed
  uint64_t freed = 0;   // listed buffers freed to make room
};
inline thread_local ReclaimList* tl_reclaim = nullptr;

// Growable device allocation, reused across builds on one handle.
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  inline void* ensure(size_t bytes);
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

// Frees one listed buffer other than `self` (after the stream has drained:
// kernels in flight may still read it) toward an allocation of `need`
// bytes: the smallest that covers what free memory lacks, else the
// largest; false if none is left.  Callers free one at a time until their
// allocation fits, so a build short of memory gives up as little of the
// next build's buffers as it can.
inline bool reclaim_dead(const DevBuf* self, size_t need) {
  ReclaimList* r = tl_reclaim;
  if (!r) return false;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
  const size_t lack = need > fr ? need - fr : 0;
  const size_t none = r->dead.size();
  size_t fit = none, big = none;
  for (size_t i = 0; i < r->dead.size(); i++) {
    const DevBuf* d = r->dead[i];
    if (d == self || !d->p) continue;
    if (d->cap >= lack && (fit == none || d->cap < r->dead[fit]->cap)) fit = i;
    if (big == none || d->cap > r->dead[big]->cap) big = i;
  }
  const size_t pick = fit != none ? fit : big;
  if (pick == none) return false;
  (void)hipStreamSynchronize(r->st);
  if (r->st2) (void)hipStreamSynchronize(r->st2);
  r->dead[pick]->release();
  r->dead[pick] = r->dead.back();
  r->dead.pop_back();
  r->freed++;
  return true;
}

// Takes a buffer in use again off the list (every listing: a buffer may be
// listed by several phases).
inline void reclaim_forget(const DevBuf* d) {
  if (!tl_reclaim) return;
  auto& v = tl_reclaim->dead;
  for (size_t i = 0; i < v.size();)
    if (v[i] == d) {
      v[i] = v.back();
      v.pop_back();
    } else {
      i++;
    }
}

inline void* DevBuf::ensure(size_t bytes) {
  reclaim_forget(this);
  if (bytes == 0) bytes = 256;
  if (bytes > cap) {
    release();
    hipError_t e = hipMalloc(&p, bytes);
    while (e == hipErrorOutOfMemory && reclaim_dead(this, bytes)) {
      (void)hipGetLastError();
      e = hipMalloc(&p, bytes);
    }
    if (tl_reclaim) tl_reclaim->allocs++;
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw HipError("hipMalloc of " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e), (int)e);
    }
    cap = bytes;
  }
  return p;
}

// Installs a handle's reclaim list for the calling thread while a build
// phase runs.
struct ReclaimScope {
  ReclaimList* prev;
  ReclaimScope(ReclaimList* r, hipStream_t st) : prev(tl_reclaim) {
    r->st = st;
    tl_reclaim = r;
  }
  ~ReclaimScope() { tl_reclaim = prev; }
  ReclaimScope(const ReclaimScope&) = delete;
  ReclaimScope& operator=(const ReclaimScope&) = delete;
};

constexpr int kBlock = 256;

static inline unsigned grid_for(size_t n, size_t per_block, unsigned cap = 1u << 20) {
  size_t g = (n + per_block - 1) / per_block;
  if (g == 0) g = 1;
  return (unsigned)(g > cap ? cap : g);
}

// ---------------------------------------------------------------------------
// Exclusive scan: out[i] = sum(in[0..i)), out[n] = total.  out may not alias in.
// ---------------------------------------------------------------------------

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <class T>
__device__ inline T block_exclusive_scan(T v, T* total) {
  __shared__ T wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  T x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; w++) {
    if (w < wid) before += wsum[w];
    all += wsum[w];
  }
  __syncthreads();
  *total = all;
  return before + x - v;
}

template <class Tin, class T>
__global__ void __launch_bounds__(kBlock) k_scan_reduce(const Tin* __restrict__ in, size_t n, T* __restrict__ part) {
  size_t base = (size_t)blockIdx.x * kScanTile;
  T s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    size_t j = base + (size_t)i * kBlock + threadIdx.x;
    if (j < n) s += (T)in[j];
  }
  T tot;
  (void)block_exclusive_scan<T>(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// The tile passes through LDS twice: loaded and stored with consecutive
// threads on consecutive elements (whole lines per wave instruction), and
// scanned with each thread on kScanItems consecutive elements.  Tile
// position q sits at LDS slot q + q / kScanItems, so the threads' runs of
// kScanItems fall on different banks.
template <class Tin, class T>
__global__ void __launch_bounds__(kBlock) k_scan_apply(const Tin* __restrict__ in, size_t n, const T* __restrict__ part,
                                                       T* __restrict__ out) {
  __shared__ T tile[kScanTile + kBlock];
  const size_t base = (size_t)blockIdx.x * kScanTile;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    const uint32_t q = (uint32_t)(i * kBlock + threadIdx.x);
    const size_t j = base + q;
    tile[q + q / kScanItems] = j < n ? (T)in[j] : (T)0;
  }
  __syncthreads();
  const uint32_t run = threadIdx.x * (kScanItems + 1);  // slot of this thread's first element
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    v[i] = tile[run + i];
    s += v[i];
  }
  T tot;
  T ex = block_exclusive_scan<T>(s, &tot) + (part ? part[blockIdx.x] : (T)0);
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {  // the thread's own slots: no other thread reads them
    tile[run + i] = ex;
    ex += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = ex;
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kScanItems; i++) {
    const uint32_t q = (uint32_t)(i * kBlock + threadIdx.x);
    const size_t j = base + q;
    if (j < n) out[j] = tile[q + q / kScanItems];
  }
}

// Scratch needed by exclusive_scan for n elements (bytes), all levels.
template <class T>
static size_t scan_scratch_bytes(size_t n) {
  size_t b = 0;
  while (n > (size_t)kScanTile) {
    n = (n + kScanTile - 1) / kScanTile;
    b += (n + 1) * sizeof(T) * 2 + 256;
  }
  return b + 256;
}

template <class Tin, class T>
static void exclusive_scan(const Tin* in, T* out, size_t n, char* scratch, hipStream_t st) {
  size_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 0) tiles = 1;
  if (tiles == 1) {
    hipLaunchKernelGGL((k_scan_apply<Tin, T>), dim3(1), dim3(kBlock), 0, st, in, n, (const T*)nullptr, out);
    GX_HIP(hipGetLastError());
    return;
  }
  T* part = reinterpret_cast<T*>(scratch);
  T* partscan = part + (tiles + 1);
  char* next = reinterpret_cast<char*>(partscan + tiles + 1);
  next = reinterpret_cast<char*>(((uintptr_t)next + 255) & ~(uintptr_t)255);
  hipLaunchKernelGGL((k_scan_reduce<Tin, T>), dim3((unsigned)tiles), dim3(kBlock), 0, st, in, n, part);
  GX_HIP(hipGetLastError());
  exclusive_scan<T, T>(part, partscan, tiles, next, st);
  hipLaunchKernelGGL((k_scan_apply<Tin, T>), dim3((unsigned)tiles), dim3(kBlock), 0, st, in, n,
                     (const T*)partscan, out);
  GX_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  Keys are up to 128 bits (KeyT =
// uint32_t, uint64_t or Key128); the digits sorted are bytes
// [byte_lo, byte_hi) of the key's integer value.  Values are uint32_t or
// uint64_t.
// ---------------------------------------------------------------------------

struct Key128 {
  uint64_t lo, hi;
};

__device__ __host__ inline uint32_t key_byte(uint32_t k, int b) { return (k >> (8 * b)) & 0xff; }
__device__ __host__ inline uint32_t key_byte(uint64_t k, int b) { return (uint32_t)(k >> (8 * b)) & 0xff; }
__device__ __host__ inline uint32_t key_byte(const Key128& k, int b) {
  return b < 8 ? (uint32_t)(k.lo >> (8 * b)) & 0xff : (uint32_t)(k.hi >> (8 * (b - 8))) & 0xff;
}

constexpr int kSortItems = 8;
constexpr int kSortTile = kBlock * kSortItems;

// Tile-local digit bases for a staged scatter: offs is the exclusive scan of
// the digit-major per-tile counts (entry d * ntiles + tile, total last), so
// this tile's count of digit d is the difference of two neighbours; lbase[d]
// = elements of the tile with a smaller digit (IdxT: any type holding
// kSortTile).  Block of kBlock = 256 threads, one digit each.
template <class IdxT>
__device__ inline void tile_digit_bases(const uint64_t* __restrict__ offs, unsigned ntiles, IdxT* lbase, IdxT* lrun,
                                        uint32_t* wsum) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const size_t at = (size_t)threadIdx.x * ntiles + blockIdx.x;
  const uint32_t c = (uint32_t)(offs[at + 1] - offs[at]);
  uint32_t incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  uint32_t before = incl - c;
  for (int w = 0; w < wid; w++) before += wsum[w];
  lbase[threadIdx.x] = (IdxT)before;
  lrun[threadIdx.x] = (IdxT)before;
}

template <class KeyT>
__global__ void __launch_bounds__(kBlock) k_radix_hist(const KeyT* __restrict__ keys, size_t n, int byte,
                                                       uint32_t* __restrict__ hist, unsigned ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int i = 0; i < kSortItems; i++) {
    size_t j = base + (size_t)i * kBlock + threadIdx.x;
    if (j < n) atomicAdd(&h[key_byte(keys[j], byte)], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Scatter with stable ranks: the tile is processed in kSortItems rounds of
// 256 consecutive elements; inside a round the rank of an element among equal
// digits comes from a wave-level match (8 ballots) plus the counts of the
// lower waves, so ties keep their input order.
// A load of data read once, in order: nontemporal for integer words, so
// the stream does not evict lines other accesses reuse.
template <class T>
__device__ inline T load_stream(const T* p) {
  if constexpr (std::is_integral<T>::value) return __builtin_nontemporal_load(p);
  else return *p;
}

// The tile's elements are ranked into an LDS stage in digit order, then
// written out as one contiguous run per digit (the tile's share of that
// digit's output range): scattered element stores would leave lines partly
// written across rounds, and the partial lines cost extra HBM writes.
template <class KeyT, bool kVals, class ValT = uint32_t>
__global__ void __launch_bounds__(kBlock) k_radix_scatter(const KeyT* __restrict__ keys, const ValT* __restrict__ vals,
                                                          size_t n, int byte, const uint64_t* __restrict__ offs,
                                                          unsigned ntiles, KeyT* __restrict__ okeys,
                                                          ValT* __restrict__ ovals) {
  __shared__ uint64_t run[256];
  __shared__ uint32_t lbase[256], lrun[256], wsum[kBlock / 64];
  __shared__ uint16_t wcnt[kBlock / 64][256];
  __shared__ KeyT sk[kSortTile];
  __shared__ ValT sv[kVals ? kSortTile : 1];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  run[threadIdx.x] = offs[(size_t)threadIdx.x * ntiles + blockIdx.x];
  tile_digit_bases(offs, ntiles, lbase, lrun, wsum);
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  size_t base = (size_t)blockIdx.x * kSortTile;
  for (int r = 0; r < kSortItems; r++) {
    for (int w = 0; w < kBlock / 64; w++) wcnt[w][threadIdx.x] = 0;
    __syncthreads();
    size_t j = base + (size_t)r * kBlock + threadIdx.x;
    bool valid = j < n;
    KeyT key{};
    ValT val = 0;
    uint32_t d = 0;
    if (valid) {
      key = load_stream(keys + j);
      if (kVals) val = load_stream(vals + j);
      d = key_byte(key, byte);
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
      sk[pos] = key;
      if (kVals) sv[pos] = val;
    }
    __syncthreads();
    uint32_t tot = 0;
    for (int w = 0; w < kBlock / 64; w++) tot += wcnt[w][threadIdx.x];
    lrun[threadIdx.x] += tot;
  }
  __syncthreads();
  const uint32_t tn = (uint32_t)(n - base < (size_t)kSortTile ? n - base : (size_t)kSortTile);
  for (uint32_t i = threadIdx.x; i < tn; i += kBlock) {
    const KeyT key = sk[i];
    const uint32_t d = key_byte(key, byte);
    const uint64_t pos = run[d] + (i - lbase[d]);
    okeys[pos] = key;
    if (kVals) ovals[pos] = sv[i];
  }
}

template <class KeyT>
static size_t radix_scratch_bytes(size_t n) {
  size_t ntiles = (n + kSortTile - 1) / kSortTile;
  if (ntiles == 0) ntiles = 1;
  size_t h = 256 * ntiles;
  return h * 4 + (h + 1) * 8 + scan_scratch_bytes<uint64_t>(h) + 1024;
}

// Sorts (keys, vals) by bytes [byte_lo, byte_hi).  Uses (k2, v2) as the
// ping-pong buffers; returns true when the result ended in (k2, v2).
template <class KeyT, class ValT = uint32_t>
static bool radix_sort(KeyT* k1, ValT* v1, KeyT* k2, ValT* v2, size_t n, int byte_lo, int byte_hi,
                       char* scratch, hipStream_t st) {
  if (n <= 1) return false;
  unsigned ntiles = (unsigned)((n + kSortTile - 1) / kSortTile);
  size_t h = (size_t)256 * ntiles;
  uint32_t* hist = reinterpret_cast<uint32_t*>(scratch);
  uint64_t* offs = reinterpret_cast<uint64_t*>(((uintptr_t)(hist + h) + 255) & ~(uintptr_t)255);
  char* sc = reinterpret_cast<char*>(((uintptr_t)(offs + h + 1) + 255) & ~(uintptr_t)255);
  bool flip = false;
  for (int b = byte_lo; b < byte_hi; b++) {
    KeyT* ik = flip ? k2 : k1;
    KeyT* ok = flip ? k1 : k2;
    ValT* iv = flip ? v2 : v1;
    ValT* ov = flip ? v1 : v2;
    hipLaunchKernelGGL((k_radix_hist<KeyT>), dim3(ntiles), dim3(kBlock), 0, st, ik, n, b, hist, ntiles);
    GX_HIP(hipGetLastError());
    exclusive_scan<uint32_t, uint64_t>(hist, offs, h, sc, st);
    if (v1)
      hipLaunchKernelGGL((k_radix_scatter<KeyT, true, ValT>), dim3(ntiles), dim3(kBlock), 0, st, ik, iv, n, b, offs,
                         ntiles, ok, ov);
    else
      hipLaunchKernelGGL((k_radix_scatter<KeyT, false, ValT>), dim3(ntiles), dim3(kBlock), 0, st, ik, iv, n, b,
                         offs, ntiles, ok, ov);
    GX_HIP(hipGetLastError());
    flip = !flip;
  }
  return flip;
}

}  // namespace gx
