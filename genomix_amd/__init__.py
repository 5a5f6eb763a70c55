"""MI355X-native Genomix graph build and unitig compression.

The hot path runs in hand-written HIP kernels for gfx950 behind a C ABI
(include/genomix_gx.h, libgenomix_gx.so); this package is the host-side
mirror of the reference's graph-build entry (GenomixHyracksDriver.runJob
with Plan.BUILD_DEBRUIJN_GRAPH) plus the SequenceFile writer around it, and
of the GenomixDriver steps on the built graph: cleaning (REMOVE_BAD_COVERAGE,
TIP_SINGLE_NODE, BUBBLE), statistics, FASTA and the coverage cutoff
(GraphBuilder methods; fit_mixture for binned data).
"""

from .gx import GraphBatch, GraphBuilder, GxError, build_graph, fit_mixture  # noqa: F401
