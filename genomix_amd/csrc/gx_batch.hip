// Batched graph builds: many read sets of uneven size and mixed k on one GPU.
//
// A native scheduler: the sets are ordered by input size, largest first
// (longest-processing-time list scheduling), and handed out through an
// atomic counter to `workers` host threads.  Each worker owns one build
// handle (its own HIP stream and device buffers, re-targeted to each set's
// k with gx_build_set_config), so the host-side steps of one set (size
// read-backs, launches) overlap the device work of the others and small
// sets fill the GPU while a large one is between phases.  Per size class
// (k-mer occurrences) the worker applies the tuned aggregation method, grid
// and learned-table slack.  Outputs are handed to the caller's consumer while the handle
// still holds them.  No counterpart in the reference (one Hyracks job per
// read set).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/genomix_batch.h"
#include "../../include/genomix_stream.h"

namespace {

struct SetSpec {
  gx_build_cfg cfg{};
  const void* input = nullptr;
  size_t len = 0;
  int fmt = GX_FMT_READIDS;
  int on_device = 0;
  bool given = false;
  uint64_t stats[9] = {};
  double ms = 0;
};

struct Tuning {
  int workers = 4;
  std::vector<uint64_t> max_occ{1ull << 23, 1ull << 27};
  std::vector<int> grid{1024, 2048, 2048};
  std::vector<float> slack{1.5f, 1.5f, 1.5f};
  std::vector<int> method{GX_AGG_SORT, GX_AGG_SORT, GX_AGG_SORT};
};

// rough k-mer occurrence count of a readids buffer (for the size class
// before the build knows it): bytes x (L - k + 1) / (L + id and separators)
uint64_t occ_estimate(const SetSpec& s) { return (uint64_t)((double)s.len * 0.8); }

}  // namespace

struct gx_batch {
  int device = 0;
  std::vector<SetSpec> sets;
  gx_mlp* sizer = nullptr;
  Tuning tune;
  std::string err;
  std::mutex mu;
  std::vector<gx_build*> pool;  // one handle per worker, kept across runs
  int retried = 0;              // sets of the last run taken again alone after running out of memory
};

extern "C" {

int gx_batch_create(int device, int nsets, gx_batch** out) {
  if (!out || nsets < 0) return GX_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return GX_EDEVICE;
  if (device < 0 || device >= ndev) return GX_EINVAL;
  gx_batch* b = new (std::nothrow) gx_batch;
  if (!b) return GX_ENOMEM;
  b->device = device;
  b->sets.resize((size_t)nsets);
  *out = b;
  return GX_OK;
}

int gx_batch_set(gx_batch* b, int set, const gx_build_cfg* cfg, const void* input, size_t len, int fmt,
                 int on_device) {
  if (!b) return GX_EINVAL;
  if (!cfg || set < 0 || set >= (int)b->sets.size() || (len && !input)) {
    b->err = "bad set index or null argument";
    return GX_EINVAL;
  }
  if (cfg->device != b->device) {
    b->err = "every set runs on the batch's device";
    return GX_EINVAL;
  }
  if (fmt != GX_FMT_READIDS && fmt != GX_FMT_FASTQ) {
    b->err = "batch inputs are readids or single-end FASTQ buffers";
    return GX_EINVAL;
  }
  SetSpec& s = b->sets[(size_t)set];
  s.cfg = *cfg;
  s.input = input;
  s.len = len;
  s.fmt = fmt;
  s.on_device = on_device != 0;
  s.given = true;
  return GX_OK;
}

int gx_batch_set_sizer(gx_batch* b, gx_mlp* m) {
  if (!b) return GX_EINVAL;
  b->sizer = m;
  return GX_OK;
}

int gx_batch_set_tuning(gx_batch* b, int workers, int nclasses, const uint64_t* max_occ, const int* grid_blocks,
                        const float* table_slack) {
  if (!b) return GX_EINVAL;
  if (workers < 1 || workers > 8 || nclasses < 1 || nclasses > 16 || !grid_blocks || !table_slack ||
      (nclasses > 1 && !max_occ)) {
    b->err = "bad tuning (1..8 workers, 1..16 classes)";
    return GX_EINVAL;
  }
  Tuning t;
  t.workers = workers;
  t.max_occ.assign(max_occ, max_occ + (nclasses - 1));
  t.grid.assign(grid_blocks, grid_blocks + nclasses);
  t.slack.assign(table_slack, table_slack + nclasses);
  t.method.assign((size_t)nclasses, GX_AGG_SORT);
  b->tune = t;
  return GX_OK;
}

int gx_batch_set_methods(gx_batch* b, int nclasses, const int* method) {
  if (!b) return GX_EINVAL;
  if (nclasses != (int)b->tune.grid.size() || !method) {
    b->err = "methods for " + std::to_string(nclasses) + " classes; the tuning has " +
             std::to_string(b->tune.grid.size());
    return GX_EINVAL;
  }
  for (int c = 0; c < nclasses; c++)
    if (method[c] != GX_AGG_SORT && method[c] != GX_AGG_HASH) {
      b->err = "unknown aggregation method for class " + std::to_string(c);
      return GX_EINVAL;
    }
  b->tune.method.assign(method, method + nclasses);
  return GX_OK;
}

int gx_batch_run(gx_batch* b, int unitigs, gx_batch_consumer consumer, void* user) {
  if (!b) return GX_EINVAL;
  const int n = (int)b->sets.size();
  for (int i = 0; i < n; i++)
    if (!b->sets[(size_t)i].given) {
      b->err = "set " + std::to_string(i) + " has no configuration";
      return GX_EINVAL;
    }
  std::vector<int> order((size_t)n);
  for (int i = 0; i < n; i++) order[(size_t)i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return b->sets[(size_t)x].len > b->sets[(size_t)y].len; });
  std::atomic<int> next{0};
  std::atomic<int> failed{0};
  int first_rc = GX_OK;
  std::vector<int> retry;  // sets that ran out of device memory beside the other workers' caches
  const int nw = std::max(1, std::min(b->tune.workers, n));
  if ((int)b->pool.size() < nw) b->pool.resize((size_t)nw, nullptr);
  auto fail = [&](int rc, const std::string& msg) {
    std::lock_guard<std::mutex> g(b->mu);
    if (first_rc == GX_OK) {
      first_rc = rc;
      b->err = msg;
    }
    failed = 1;
  };
  // One set on handle h (created on first use).  Returns GX_OK, or the
  // set's error code with *msg set.
  auto run_set = [&](gx_build*& h, int si, std::string* msg) -> int {
    SetSpec& s = b->sets[(size_t)si];
    const std::string where = "set " + std::to_string(si) + ": ";
    int rc = GX_OK;
    if (!h) {
      std::lock_guard<std::mutex> g(b->mu);  // creation errors are reported through one global string
      rc = gx_build_create(&s.cfg, &h);
      if (rc) {
        *msg = where + gx_last_error(nullptr);
        h = nullptr;
        return rc;
      }
    } else {
      rc = gx_build_set_config(h, &s.cfg);
    }
    size_t c = 0;
    uint64_t est = occ_estimate(s);
    while (c < b->tune.max_occ.size() && est > b->tune.max_occ[c]) c++;
    // one pass per set: the free-memory estimate behind automatic key
    // ranges would count the other workers' buffers against this one
    if (!rc && unitigs) rc = gx_build_set_stream(h, ~0ull, nullptr, nullptr);
    if (!rc) rc = gx_build_set_tuning(h, b->tune.grid[c], b->tune.slack[c]);
    if (!rc) rc = gx_build_set_method(h, b->tune.method[c]);
    if (!rc) rc = gx_build_set_sizer(h, b->sizer);
    auto t0 = std::chrono::steady_clock::now();
    if (!rc)
      rc = s.on_device ? gx_build_add_reads_device(h, s.input, s.len, s.fmt)
                       : gx_build_add_reads(h, (const char*)s.input, s.len, s.fmt);
    if (!rc) rc = gx_build_finish(h);
    if (!rc && unitigs) rc = gx_unitig_run(h);
    auto t1 = std::chrono::steady_clock::now();
    if (rc) {
      *msg = where + gx_last_error(h);
      return rc;
    }
    s.ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    gx_build_stats(h, s.stats, 9);
    if (consumer) {
      int crc = consumer(user, si, h);
      if (crc) {
        *msg = where + "consumer returned " + std::to_string(crc);
        return crc < 0 ? crc : GX_EINVAL;
      }
    }
    return GX_OK;
  };
  auto worker = [&](int w) {
    gx_build*& h = b->pool[(size_t)w];
    if (hipSetDevice(b->device) != hipSuccess) {
      fail(GX_EDEVICE, "hipSetDevice failed");
      return;
    }
    for (;;) {
      if (failed) break;
      int j = next.fetch_add(1);
      if (j >= n) break;
      const int si = order[(size_t)j];
      std::string msg;
      const int rc = run_set(h, si, &msg);
      if (rc == GX_ENOMEM) {
        // the device memory is held by the workers' buffers: give this
        // worker's back and run the set again alone at the end
        if (h) gx_build_destroy(h);
        h = nullptr;
        std::lock_guard<std::mutex> g(b->mu);
        retry.push_back(si);
        continue;
      }
      if (rc) {
        fail(rc, msg);
        break;
      }
    }
  };
  std::vector<std::thread> threads;
  for (int w = 0; w < nw; w++) threads.emplace_back(worker, w);
  for (auto& t : threads) t.join();
  if (!failed && !retry.empty()) {
    for (gx_build*& h : b->pool)
      if (h) {
        gx_build_destroy(h);
        h = nullptr;
      }
    std::sort(retry.begin(), retry.end());
    if (hipSetDevice(b->device) != hipSuccess) return GX_EDEVICE;
    for (int si : retry) {
      std::string msg;
      const int rc = run_set(b->pool[0], si, &msg);
      if (rc) {
        fail(rc, msg);
        break;
      }
    }
  }
  b->retried = (int)retry.size();
  return first_rc;
}

int gx_batch_set_stats(gx_batch* b, int set, uint64_t* out, int n, double* ms) {
  if (!b || set < 0 || set >= (int)b->sets.size()) return GX_EINVAL;
  const SetSpec& s = b->sets[(size_t)set];
  if (out)
    for (int i = 0; i < n && i < 9; i++) out[i] = s.stats[i];
  if (ms) *ms = s.ms;
  return GX_OK;
}

int gx_batch_retried(const gx_batch* b) { return b ? b->retried : 0; }

const char* gx_batch_last_error(const gx_batch* b) { return b ? b->err.c_str() : "null handle"; }

void gx_batch_destroy(gx_batch* b) {
  if (!b) return;
  for (gx_build* h : b->pool)
    if (h) gx_build_destroy(h);
  delete b;
}

}  // extern "C"
