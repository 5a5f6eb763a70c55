"""Host-side Node-file codec (product) against the oracle's independent one
and the reference's binary SequenceFile fixtures."""

import glob
import os

import pytest

from genomix_amd import formats
from oracle import oracle

from .conftest import GOLDEN

SEQFILES = sorted(glob.glob(os.path.join(GOLDEN, "seqfile", "*", "part-*")))


@pytest.mark.parametrize("path", SEQFILES, ids=[os.path.relpath(p, GOLDEN) for p in SEQFILES])
def test_seqfile_round_trip_on_reference_fixtures(path):
    data = open(path, "rb").read()
    kc, vc, sync, recs = formats.read_seqfile(data)
    assert formats.seqfile_bytes(recs, sync, kc, vc) == data
    crc = os.path.join(os.path.dirname(path), "." + os.path.basename(path) + ".crc")
    if os.path.exists(crc):
        assert formats.crc_sidecar(data) == open(crc, "rb").read()


def test_text_dump_matches_oracle():
    for exp in sorted(glob.glob(os.path.join(GOLDEN, "hyracks", "**", "*.test.txt"), recursive=True)):
        text = open(exp, "rb").read()
        recs = oracle.build_graph(text, 3)[0]
        assert formats.text_dump(recs) == oracle.dump_text(recs)
        u = oracle.unitigs(recs, 3)
        assert formats.text_dump(u) == oracle.dump_text(u)
        sync = bytes(range(16))
        assert formats.seqfile_bytes(recs, sync) == oracle.seqfile(recs, sync)


@pytest.mark.parametrize("f", [1.0, 8.0, 445.0, 0.1, 1.0 / 3.0, 12345678.0, 2.5, 37.333332])
def test_java_float(f):
    assert formats.java_float(f) == oracle.float_str(f)
