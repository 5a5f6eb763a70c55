/* genomix_batch.h — many read sets, mixed k, one call.
 *
 * No counterpart in the reference: Genomix runs one Hyracks job per read
 * set.  These entry points are for callers that build many sets at once
 * (one pipeline run over a sample collection): a native scheduler hands the
 * sets, largest first, to worker threads, each driving one build handle on
 * its own HIP stream of the same GPU, so small sets and host-side steps of
 * one set overlap device work of the others.  Each set's output is exactly
 * what a separate gx_build handle would produce for it.
 *
 * All functions return GX_OK (0) or a negative GX_* code (genomix_gx.h).
 */
#ifndef GENOMIX_BATCH_H
#define GENOMIX_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "genomix_gx.h"
#include "genomix_sizer.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-handle launch and sizing knobs (defaults: 2048 blocks for grid-stride
 * kernels, learned table = 1.5 x predicted distinct). */
int gx_build_set_tuning(gx_build* b, int grid_blocks, float table_slack);

/* Re-targets an existing handle to another configuration (k, nparts, edge
 * order; same device), keeping its device buffers: a handle reused across
 * read sets of different k. Clears any input added so far. */
int gx_build_set_config(gx_build* b, const gx_build_cfg* cfg);

/* Aggregation method of a handle.  GX_AGG_SORT (default): occurrences
 * sorted by (key, occurrence) -- radix passes on the top key bits, then
 * per-bucket sorts in LDS -- and nodes taken as runs; streaming HBM traffic
 * only.  GX_AGG_HASH: an open-addressing table with atomic count/mask
 * updates, sized up front (learned sizing applies here) and rebuilt on
 * overflow.  Output bytes are identical. */
#define GX_AGG_SORT 0
#define GX_AGG_HASH 1
int gx_build_set_method(gx_build* b, int method);

typedef struct gx_batch gx_batch;

/* Called on a worker thread once set `set` is built (and merged, if asked):
 * the handle holds its records until the callback returns.  Callbacks of
 * different workers run concurrently.  A non-zero return aborts the batch. */
typedef int (*gx_batch_consumer)(void* user, int set, gx_build* handle);

int gx_batch_create(int device, int nsets, gx_batch** out);
/* Configuration and input of one set.  The input stays owned by the caller
 * and must stay valid until gx_batch_run returns (on_device = 1: a device
 * pointer on the batch's device). */
int gx_batch_set(gx_batch* b, int set, const gx_build_cfg* cfg, const void* input, size_t len, int fmt,
                 int on_device);
int gx_batch_set_sizer(gx_batch* b, gx_mlp* m);
/* Size classes by k-mer occurrences: class c holds sets with at most
 * max_occ[c] occurrences (the last class takes the rest); each class has
 * its grid and table slack.  workers = worker threads / streams (1..8). */
int gx_batch_set_tuning(gx_batch* b, int workers, int nclasses, const uint64_t* max_occ, const int* grid_blocks,
                        const float* table_slack);
/* Aggregation method per size class (GX_AGG_SORT or GX_AGG_HASH; nclasses
 * as in the last gx_batch_set_tuning, which resets every class to
 * GX_AGG_SORT).  With GX_AGG_HASH and a sizing network attached, the
 * network sizes the class's tables (the slack above). */
int gx_batch_set_methods(gx_batch* b, int nclasses, const int* method);
/* Builds every set (and runs unitig compression when unitigs != 0),
 * calling consumer (may be NULL) after each.  With unitigs each set is
 * built in one pass (the free-memory estimate behind key ranges would count
 * the other workers' buffers); a set that runs out of device memory beside
 * those buffers is built again alone after the others. */
int gx_batch_run(gx_batch* b, int unitigs, gx_batch_consumer consumer, void* user);
/* Sets of the last run that were built again alone. */
int gx_batch_retried(const gx_batch* b);
/* Per-set statistics of the last run (gx_build_stats layout, n <= 9) and
 * its device time in milliseconds (finish + unitigs, host clock). */
int gx_batch_set_stats(gx_batch* b, int set, uint64_t* out, int n, double* ms);
const char* gx_batch_last_error(const gx_batch* b);
void gx_batch_destroy(gx_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* GENOMIX_BATCH_H */
