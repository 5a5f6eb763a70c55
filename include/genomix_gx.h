/* genomix_gx.h — C ABI of the MI355X graph build and unitig compression.
 *
 * Drop-in boundary for the Genomix Hyracks graph-build path.  In the
 * reference, GenomixHyracksDriver.runJob (genomix-hyracks/.../graph/driver/
 * GenomixHyracksDriver.java:84-135) builds a Hyracks job whose operators are
 *   HDFSReadOperatorDescriptor -> ReadsKeyValueParserFactory (IKeyValueParser)
 *   -> ExternalSort(KmerPointable) -> PreclusteredGroup(AggregateKmerAggregate)
 *   -> MToN hash partition -> PreclusteredGroup(AggregateKmerAggregate)
 *   -> HDFSWriteOperatorDescriptor(KmerNodePairSequenceWriterFactory)
 * (JobGenBuildBrujinGraph.java:79-154).  A gx_build handle replaces the
 * parser -> sort -> local group -> MToN -> cross group chain of ONE Hyracks
 * task: reads go in as the same readids text lines the parser consumes, and
 * Node records come out in the exact byte form the SequenceFile writer
 * appends.  All entry points return 0 on success and a negative GX_E* code
 * on failure (gx_last_error gives the message); nothing throws across the
 * ABI.  A handle is not thread-safe; distinct handles may run concurrently,
 * each on its own HIP streams (a main stream, and a second one created on
 * first use for work that overlaps it; every call has finished its work on
 * both when it returns).  The library owns all device memory it allocates;
 * callers own the buffers they pass in.
 */
#ifndef GENOMIX_GX_H
#define GENOMIX_GX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define GX_OK 0
#define GX_EINVAL -22   /* bad argument or configuration (HyracksDataException / IllegalArgumentException) */
#define GX_EINPUT -5    /* malformed input line; the reference job fails on the same line */
#define GX_ENOMEM -12   /* device allocation failed */
#define GX_EDEVICE -19  /* no usable HIP device / HIP runtime error */
#define GX_ESTATE -16   /* call out of order (e.g. emit before finish) */
#define GX_ENOTSUP -95  /* option not implemented in this build */

/* input formats */
#define GX_FMT_READIDS 0 /* "readId\tseq[\tmateSeq]" lines (library-N.readids) */
#define GX_FMT_FASTQ 1   /* single-end FASTQ of one library; converted on the device as
                            GenomixDriver.convertAndUploadFastqToHDFS does (GenomixDriver.java:665-714) */

/* edge-list order inside each FF/FR/RF/RR list */
#define GX_EDGE_ORDER_CANONICAL 0 /* ascending neighbour k-mer (the reference's tests compare lists as sets) */
#define GX_EDGE_ORDER_JDK7 1      /* byte-identical to the reference at P=1: JDK-7 HashSet replay */

typedef struct gx_build gx_build;

typedef struct {
  int k;                    /* GenomixJobConf KMER_LENGTH (GenomixJobConf.java:303); 3 <= k <= 63 */
  int nparts;               /* output partitions P, 1..65536 (Hyracks partitions; 1 = CLI default).  A record
                               goes to part nonneg(hashBytes(key)) % P (KmerPartitionComputerFactory.java:28-53),
                               each part in KmerPointable order: the reference's part-<p> for one input split */
  int max_readids_per_edge; /* GenomixJobConf MAX_READIDS_PER_EDGE (:333); validated >= 0, no effect (as in the reference) */
  int edge_order;           /* GX_EDGE_ORDER_* */
  int device;               /* HIP device ordinal */
  size_t table_hint;        /* expected distinct k-mers (0 = size from the input) */
} gx_build_cfg;

/* Replaces the per-task set-up of ReadsKeyValueParserFactory.createKeyValueParser
 * (ReadsKeyValueParserFactory.java:64-80), AggregateKmerAggregateFactory.createAggregator
 * (AggregateKmerAggregateFactory.java:54-66) and GenomixJobConf.setGlobalStaticConstants
 * (GenomixJobConf.java:689-694): validates k (GenomixJobConf.java:417-423) and binds a device. */
int gx_build_create(const gx_build_cfg* cfg, gx_build** out);

/* Replaces IKeyValueParser.parse (hyracks-hdfs/.../api/IKeyValueParser.java:29-57;
 * ReadsKeyValueParserFactory.java:95-148) for a whole buffer of lines: the
 * text is copied (host -> HBM) before return.  Buffers concatenate, so a line
 * may span two calls. */
int gx_build_add_reads(gx_build* b, const char* buf, size_t len, int fmt);

/* Same, for text already resident in device memory (no host copy; the
 * buffer is copied device -> device into the handle before return). */
int gx_build_add_reads_device(gx_build* b, const void* dev_buf, size_t len, int fmt);

/* Paired-end FASTQ of one library: mate-1 and mate-2 file bytes, each
 * appended to its own stream (chunks concatenate per file).  At finish the
 * two files are converted on the device into "readId\tmate1\tmate2" lines,
 * readId = 1-based line number of each sequence line (4i+2), both sequences
 * String.trim()ed, as GenomixDriver.convertAndUploadFastqToHDFS
 * (GenomixDriver.java:665-714); files with different line counts fail with
 * GX_EINPUT ("didn't have the same number of lines").  A handle takes one
 * input format: mixing formats across add calls is GX_EINVAL. */
int gx_build_add_fastq_pair(gx_build* b, const char* mate1, size_t n1, const char* mate2, size_t n2);

/* After finish on FASTQ input: the library-N.readids text the device made
 * (the file convertAndUploadFastqToHDFS writes, GenomixDriver.java:669-712).
 * dst == NULL stores the size in *written. */
int gx_build_readids(gx_build* b, char* dst, size_t cap, size_t* written);

/* Runs the whole build on the device (split reads, canonical k-mers, table
 * insert, aggregation, read heads, record serialisation) and waits for it.
 * Replaces ExternalSort + both PreclusteredGroup operators + the MToN
 * connector (JobGenBuildBrujinGraph.java:92-117). */
int gx_build_finish(gx_build* b);

/* Number of records and bytes of one output partition (part-<p>). */
int gx_build_part_info(gx_build* b, int part, uint64_t* nrec, uint64_t* nbytes);

/* Copies partition `part` into dst: SequenceFile records
 * (i32 recLen, i32 keyLen, VKmer key, Node value) in KmerPointable order,
 * without the file header and sync markers, i.e. the bytes
 * KmerNodePairSequenceWriterFactory.write (KmerNodePairSequenceWriterFactory.java:79-94)
 * hands to SequenceFile.Writer.append. */
int gx_build_emit_part(gx_build* b, int part, uint8_t* dst, size_t cap, size_t* written);

/* Device pointer to partition `part`'s record bytes (valid until the next
 * finish/reset/destroy on this handle). */
int gx_build_part_device(gx_build* b, int part, const void** dev_ptr, uint64_t* nbytes);

/* Unitig compression of the built graph (Genomix MERGE step: P4 path merge,
 * P4ForPathMergeVertex run by Pregelix through
 * DeBruijnGraphCleanVertex.getConfiguredJob, DeBruijnGraphCleanVertex.java:188-206,
 * with NodeToVertexInputFormat / VertexToNodeOutputFormat around it).  Every
 * maximal path whose links have degree 1 on both sides (no tandem repeats)
 * becomes one Node record: key = the path's first k-mer in KmerPointable
 * order (P4 keeps a random one), internal k-mer = the path sequence oriented
 * so that key reads forward (dropped for one-k-mer paths, as
 * GenericVertexToNodeOutputFormat.java:44-47 does), coverage = mean k-mer
 * count, read-head offsets moved to unitig positions, outer edges re-targeted
 * to neighbour unitigs (each list ascending).  Records in key order.  Run
 * after gx_build_finish; the build's own records stay available. */
int gx_unitig_run(gx_build* b);
int gx_unitig_info(gx_build* b, uint64_t* nrec, uint64_t* nbytes);
int gx_unitig_emit(gx_build* b, uint8_t* dst, size_t cap, size_t* written);

/* How the unitig pass ranks the nodes of its merge paths:
 * GX_RANK_AUTO (default) walks each path from both its endpoints toward
 * each other when k is odd and falls back to pointer jumping for even k
 * (palindromic k-mers), cycles and walks of more than 4096 nodes (every
 * path of more than 8192); GX_RANK_POINTER_JUMPING always uses pointer
 * jumping.  The records are the same either way. */
#define GX_RANK_AUTO 0
#define GX_RANK_POINTER_JUMPING 1
int gx_unitig_set_ranking(gx_build* b, int mode);

/* Limits of the unitig pass, for callers that plan around them (no device
 * needed): max_nodes = the distinct k-mers one unitig pass, cleaning step or
 * BUBBLE accepts (2^30; larger graphs fail with GX_ENOTSUP);
 * max_head_offset = the last unitig letter a read head may sit at
 * (2^23 - 1, the offset field of ReadHeadInfo.makeUUID,
 * ReadHeadInfo.java:115-119, which throws beyond it; here the unitig pass
 * fails with GX_ENOTSUP).  Either pointer may be NULL. */
int gx_unitig_limits(uint64_t* max_nodes, uint64_t* max_head_offset);

/* Statistics of the last finish: [0] occurrences, [1] distinct k-mers,
 * [2] valid mates, [3] table capacity, [4] replayed nodes, [5] replayed
 * occurrences, [6] largest coverage above 255 (0 if none), [7] distinct
 * count predicted by an attached sizing network (0 if none, genomix_sizer.h),
 * [8] table rebuilds after overflow, [9] records sorted by the global
 * fallback for oversized key buckets (sorted aggregation), [10] pointer-
 * jumping rounds of the last unitig run, [11] key-range passes of the build
 * (genomix_stream.h), [12] successor records read by those pointer-jumping
 * rounds (node ends still following their chain, summed over the rounds),
 * [13] path endpoints that walked their paths in the last unitig run (0 when
 * it ranked by pointer jumping only), [14] device allocations the last
 * finish made (0 when a repeated build reuses every buffer of the handle),
 * [15] buffers of earlier builds or phases it freed to make room. */
int gx_build_stats(gx_build* b, uint64_t* out, int n);

/* Per-phase device timings of the last finish, measured with hipEvents on
 * the handle's stream (the C-ABI counterpart of GenomixJobConf.tick/tock,
 * GenomixJobConf.java:669-687).  set_profiling(1) enables recording;
 * phase_times fills ms[0..n) and returns the number of phases;
 * phase_name(i) names phase i.  Phases named after a kernel ("k_insert":
 * the k-mer pass, "k_occ_scatter": its scatter launch alone, "k_replay",
 * "k_pj_round": the pointer-jumping launches) time those launches only;
 * the others are stages of the build or of the unitig run. */
int gx_build_set_profiling(gx_build* b, int on);
int gx_build_phase_times(gx_build* b, double* ms, int n);
const char* gx_build_phase_name(int i);

/* Drops the reads and results but keeps device allocations for reuse. */
int gx_build_reset(gx_build* b);

/* Message of the last failing call on this handle; with b NULL, of the
 * calling thread's last failing call that takes no handle (gx_build_create,
 * gx_fit_mixture). */
const char* gx_last_error(const gx_build* b);

void gx_build_destroy(gx_build* b);

#ifdef __cplusplus
}
#endif
#endif /* GENOMIX_GX_H */
