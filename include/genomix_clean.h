/* genomix_clean.h — graph cleaning on the built graph.
 *
 * Genomix's default pipeline after BUILD and MERGE runs REMOVE_BAD_COVERAGE,
 * TIP_SINGLE_NODE and BUBBLE, each followed by MERGE
 * (GenomixJobConf.java:511-517).  Each entry point here performs one step
 * and its merge on the graph a handle holds.
 *
 * All functions return GX_OK (0) or a negative GX_* code (genomix_gx.h).
 */
#ifndef GENOMIX_CLEAN_H
#define GENOMIX_CLEAN_H

#include "genomix_gx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* REMOVE_BAD_COVERAGE, then MERGE.  Replaces the Pregelix jobs
 * RemoveBadCoverageVertex (RemoveBadCoverageVertex.java:38-70: a merged
 * vertex whose average coverage is below min_cov or above max_cov is
 * removed and its neighbours drop their edges to it; conf keys
 * genomix.removeBadCoverage.minCoverage / maxCoverage, defaults 3.0 and
 * +inf, GenomixJobConf.java:330-331,498-502) and the P4 merge that follows.
 * Call after gx_unitig_run.  Afterwards gx_unitig_info / gx_unitig_emit
 * give the merged records of the cleaned graph, and the handle's graph is
 * the cleaned node graph (a further gx_unitig_run merges it again); the
 * build's own records (gx_build_emit_part) are unchanged.  Negative bounds
 * fail with GX_EINVAL, as the reference's configuration check does. */
int gx_clean_bad_coverage(gx_build* b, float min_cov, float max_cov);

/* TIP_SINGLE_NODE, then MERGE.  Replaces SingleNodeTipRemoveVertex
 * (SingleNodeTipRemoveVertex.java:47-85): a merged vertex with exactly one
 * edge on one side and none on the other (Node.degree, Node.java:820-828)
 * whose sequence has at most max_len letters is removed and its neighbour
 * drops the edge to it; then the P4 merge.  max_len < 0 means the
 * reference's default, k (genomix.tipRemove.maxLength,
 * GenomixJobConf.java:332,504-505).  Same calling rules as
 * gx_clean_bad_coverage. */
int gx_clean_tips(gx_build* b, int max_len);

/* BUBBLE, then MERGE.  Replaces SimpleBubbleMergeVertex
 * (SimpleBubbleMergeVertex.java:51-320): a merged vertex with one edge on
 * each side, to two different vertices, and fewer than max_len letters is
 * a bubble; bubbles between the same two vertices (attached to the same
 * sides of both) are taken in decreasing coverage, and each later one whose
 * edit distance to the first, over the shorter length, is at most
 * max_dissimilarity (against its reverse complement when stored the other
 * way round) is removed and merged into it: its coverage added as k-mer
 * counts, its read heads at offsets scaled to the kept one's length
 * (Node.addFromNode, Node.java:582-614; ReadHeadSet.unionUpdate,
 * ReadHeadSet.java:80-101).  Then the P4 merge.  Defaults
 * (genomix.bubbleMerge.maxDissimilarity / maxLength,
 * GenomixJobConf.java:478-482): 0.05 and 5k; max_len < 0 means 5k.
 * max_dissimilarity outside [0, 1] fails with GX_EINVAL (:429-432);
 * max_len above 4097 with GX_ENOTSUP.  Ties in coverage are broken by
 * the order of the merged records (KmerPointable order of their keys; the
 * reference's order there is its message arrival order).  Same calling rules as gx_clean_bad_coverage. */
int gx_clean_bubbles(gx_build* b, float max_dissimilarity, int max_len);

#ifdef __cplusplus
}
#endif

#endif /* GENOMIX_CLEAN_H */
