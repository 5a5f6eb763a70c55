/* genomix_stream.h — graph builds larger than one pass fits in HBM.
 *
 * The reference sorts all k-mer occurrences with Hyracks' external sort,
 * spilling sorted runs to disk when its frame budget is exceeded, and
 * merges them before the group-by (JobGenBuildBrujinGraph.java:96-98,
 * ExternalSortOperatorDescriptor; FrameSorterMergeSort.java:153-185).
 * Here a build whose occurrence records would not fit in HBM runs as
 * several passes over disjoint ranges of the sort key's top byte
 * (GX_AGG_SORT only): each pass re-extracts the k-mers of every resident
 * read, keeps the records of its range, and builds, serialises and hands on
 * that range's nodes.  Every node lies in exactly one range and the ranges
 * follow KmerPointable order, so the bytes of each partition are the same
 * as a single-pass build's.
 *
 * All functions return GX_OK (0) or a negative GX_* code (genomix_gx.h).
 */
#ifndef GENOMIX_STREAM_H
#define GENOMIX_STREAM_H

#include <stddef.h>
#include <stdint.h>

#include "genomix_gx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Receives one key range's records of partition `part`: `nrec` SequenceFile
 * records (the layout gx_build_emit_part gives), `nbytes` bytes at the
 * DEVICE address `dev_bytes`, valid until the callback returns.  Calls come
 * from the thread running gx_build_finish, ranges in key order, so the
 * concatenation of a partition's calls is that partition's file body.  A
 * non-zero return aborts the build with that code (GX_EINVAL if positive). */
typedef int (*gx_part_sink)(void* user, int part, const void* dev_bytes, uint64_t nbytes, uint64_t nrec);

/* Key-range passes and record delivery of a handle.
 *  range_occurrences: k-mer occurrences per pass; ranges are cut at values
 *    of the top key byte, so one may hold more when a single value does.
 *    0 (default): one pass when it fits in the GPU's free HBM, else as few
 *    passes of even size as fit beside what the passes keep (records and
 *    node graph, when there is no sink).  The first plan takes 0.6 distinct
 *    k-mers per occurrence, or an attached sizing network's prediction
 *    (genomix_sizer.h); the ranges after each pass are planned again from
 *    the nodes and bytes the passes so far produced and the HBM free then.
 *  sink: NULL (default) keeps the records in HBM as usual (several ranges
 *    are assembled per partition after the last); otherwise each range's
 *    records go to the sink and none are kept: gx_build_part_info still
 *    reports totals, gx_build_emit_part / gx_build_part_device fail with
 *    GX_ESTATE.
 * Without a sink the ranges' nodes and read heads are also kept, in key
 * order, and a node index over them is built after the last range, so the
 * unitig pass, cleaning and statistics take a build of several ranges as
 * they take a one-pass build.  After a build with a sink they fail with
 * GX_ENOTSUP (its graph was never whole in HBM), as after a build of 2^32
 * or more distinct k-mers (node numbers are 32-bit).  The range count of
 * the last build is gx_build_stats entry 11. */
int gx_build_set_stream(gx_build* b, uint64_t range_occurrences, gx_part_sink sink, void* user);

#ifdef __cplusplus
}
#endif

#endif /* GENOMIX_STREAM_H */
