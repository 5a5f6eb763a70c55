/* genomix_stats.h — graph statistics and FASTA dump.
 *
 * The statistics are the counters of the reference's GraphStatistics
 * mapper (GraphStatistics.java:79-146), computed on the device over the
 * graph a handle holds: for each vertex its degree, sequence length,
 * rounded average coverage, read-head counts and seed score, binned,
 * summed and maximised, plus the node, path-node, tip and self-edge
 * totals.  The FASTA dump is DriverUtils.dumpGraph (DriverUtils.java:85-128).
 * The coverage cutoff is GenomixDriver's FittingMixture step over the coverage
 * counters.
 *
 * All functions return GX_OK (0) or a negative GX_* code (genomix_gx.h).
 */
#ifndef GENOMIX_STATS_H
#define GENOMIX_STATS_H

#include <stddef.h>
#include <stdint.h>

#include "genomix_gx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Which graph: the handle's node graph (the build's, or the cleaned one
 * after genomix_clean.h) or the merged vertices of the last unitig pass. */
#define GX_GRAPH_NODES 0
#define GX_GRAPH_UNITIGS 1

/* Statistic ids, the reference's counter names: "degree", "kmerLength",
 * "coverage", "unflippedReadIds", "flippedReadIds", "scaffoldSeedScore",
 * "kmerLength-with-FORWARD", "coverage-with-FORWARD",
 * "scaffoldSeedScore-with-FORWARD", and the three -with-REVERSE. */
#define GX_STAT_DEGREE 0
#define GX_STAT_KMER_LENGTH 1
#define GX_STAT_COVERAGE 2
#define GX_STAT_UNFLIPPED 3
#define GX_STAT_FLIPPED 4
#define GX_STAT_SEED_SCORE 5
#define GX_STAT_KMER_LENGTH_FORWARD 6
#define GX_STAT_COVERAGE_FORWARD 7
#define GX_STAT_SEED_SCORE_FORWARD 8
#define GX_STAT_KMER_LENGTH_REVERSE 9
#define GX_STAT_COVERAGE_REVERSE 10
#define GX_STAT_SEED_SCORE_REVERSE 11
#define GX_NUM_STATS 12

/* Computes every statistic of `graph` and keeps them in the handle. */
int gx_graph_stats(gx_build* b, int graph);
/* "totals" counters: [0] nodes, [1] pathNode, [2] tips-FORWARD,
 * [3] tips-REVERSE, [4] tips-BOTH, [5] tips-ONE, [6..9] selfEdge-FF, -FR,
 * -RF, -RR. */
int gx_graph_stat_totals(gx_build* b, uint64_t* out, int n);
/* One statistic: its "totals" counter (sum of values), its "maximum"
 * counter (at least 0, as the reference's counter starts there) and the
 * number of distinct values (bins). */
int gx_graph_stat(gx_build* b, int stat, int64_t* total, int64_t* maximum, uint64_t* nbins);
/* The "<name>-bins" counters: distinct values ascending and their counts. */
int gx_graph_stat_bins(gx_build* b, int stat, int64_t* values, uint64_t* counts, size_t cap);

/* The coverage cutoff of GenomixDriver.setCutoffCoverageByFittingMixture
 * (GenomixDriver.java:119-136; run after BUILD_HYRACKS when
 * genomix.setCutoffCoverage is true, its result becoming
 * genomix.removeBadCoverage.minCoverage when positive).  Replaces
 * FittingMixture.fittingMixture (FittingMixture.java:89-219): an EM fit of
 * an exponential (erroneous k-mers) plus a normal (genomic k-mers) mixture
 * to the coverage values -- initial exponential mean 5, normal mean 20 and
 * sd 5, priors 1/2, memberships floored at 1e-9, `iterations` M+E rounds
 * (GenomixDriver uses 10) -- and the first integer coverage c in [1, max) at
 * which prior * density of the normal exceeds that of the exponential, or 0
 * when there is none.  Densities are commons-math3 3.2's.
 * out[GX_MIXTURE_OUT]: [0] cutoff, [1] exponential mean, [2] normal mean,
 * [3] normal sd, [4] exponential prior, [5] normal prior. */
#define GX_MIXTURE_OUT 6

/* Over the "coverage-bins" counters of the last gx_graph_stats (GenomixDriver
 * fits the build's node graph, GX_GRAPH_NODES), max being the largest
 * coverage value (GraphStatistics.getCoverageStats / getMaxCoverage,
 * GraphStatistics.java:165-199).  GX_ESTATE without gx_graph_stats or with
 * no coverage counted (GenomixDriver's "No information for coverage!"). */
int gx_coverage_cutoff(gx_build* b, int iterations, double* out);

/* The same fit over n distinct values with their counts (the reference
 * expands counter `value` into `count` copies of it) and a given max.  Host
 * arithmetic only.  GX_EINVAL for no data, max 0, iterations < 0 or a
 * non-positive exponential mean (commons-math's
 * NotStrictlyPositiveException); the message is gx_last_error(NULL). */
int gx_fit_mixture(const double* values, const uint64_t* counts, size_t n, double max, int iterations,
                   double* out);

/* ">node_<key>\n<sequence>\n" for every vertex of `graph`, in key order
 * (the sequence is the merged vertex's internal k-mer, or its key).  With
 * dst NULL only *written (the byte count) is set. */
int gx_dump_fasta(gx_build* b, int graph, uint8_t* dst, size_t cap, size_t* written);

#ifdef __cplusplus
}
#endif

#endif /* GENOMIX_STATS_H */
