"""MI355X-native Genomix graph build and unitig compression.

The hot path runs in hand-written HIP kernels for gfx950 behind a C ABI
(include/genomix_gx.h, libgenomix_gx.so); this package is the host-side
mirror of the reference's graph-build entry (GenomixHyracksDriver.runJob
with Plan.BUILD_DEBRUIJN_GRAPH) plus the SequenceFile writer around it.
"""

from .gx import GraphBuilder, GxError, build_graph  # noqa: F401
