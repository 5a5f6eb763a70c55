"""Oracle primitives against reference test vectors, Java known answers and
brute force."""

import glob
import os
import random
import struct

import pytest

from oracle import oracle

from .conftest import GOLDEN

COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def rc(s):
    return "".join(COMP[c] for c in reversed(s))


def test_kmer_fixed_vectors():
    # KmerFixedTest.java:27-63
    arr = "AATAGAAG"
    assert oracle.kmer_str(oracle.pack(arr[0:], 7), 7) == "AATAGAA"
    assert oracle.kmer_str(oracle.pack(arr[1:], 7), 7) == "ATAGAAG"
    shifted = oracle.shift(oracle.pack(arr, 7), 7, arr[-1])
    assert oracle.kmer_str(shifted, 7) == "ATAGAAG"
    assert oracle.kmer_str(oracle.pack(arr[1:], 7, reverse=True), 7) == "CTTCTAT"
    # KmerFixedTest.java:66-77
    assert oracle.kmer_str(oracle.pack("AGCTGACCG", 9), 9) == "AGCTGACCG"
    # KmerFixedTest.java:80-100: prefixes of AGCTGACCGT for k = 3..10
    for k in range(3, 11):
        assert oracle.kmer_str(oracle.pack("AGCTGACCGT", k), k) == "AGCTGACCGT"[:k]


def test_packing_brute_force():
    rnd = random.Random(20261202)
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    for k in list(range(3, 64)) + [64, 100]:
        s = "".join(rnd.choice("ACGT") for _ in range(k + 5))
        w = s[:k]
        b = oracle.pack(w, k)
        nb = (k + 3) // 4
        # closed form: BE bytes of V = sum code(s_i) * 4^i
        v = sum(code[c] << (2 * i) for i, c in enumerate(w))
        assert b == v.to_bytes(nb, "big")
        assert oracle.kmer_str(b, k) == w
        assert oracle.kmer_str(oracle.pack(w, k, reverse=True), k) == rc(w)
        # rolling shift equals packing the next window
        cur = b
        for j in range(1, 6):
            cur = oracle.shift(cur, k, s[k + j - 1])
            assert cur == oracle.pack(s[j:j + k], k)


def test_lowercase_and_invalid_codes():
    assert oracle.kmer_str(oracle.pack("acgtn", 5), 5) == "ACGTA"


def test_java_hash_known_answers():
    # Arrays.hashCode(byte[]) known answers: {} -> 1, {1,2,3} -> 30817,
    # bytes of "abc" -> 126145.  VKmer.hashCode hashes the 4-byte BE header
    # plus the packed bytes; "A" packs to one zero byte after header 0,0,0,1.
    def h(bs):
        v = 1
        for x in bs:
            v = (v * 31 + (x - 256 if x > 127 else x)) & 0xFFFFFFFF
        return v - (1 << 32) if v >= 1 << 31 else v
    assert h(b"") == 1 and h(bytes([1, 2, 3])) == 30817 and h(b"abc") == 126145
    assert oracle.vkmer_hash("A") == h(bytes([0, 0, 0, 1, 0]))
    for s in ("ACG", "TTTT", "GATTACA", "ACGTACGTACGTACGTACGTACGTACGTACGTACG"):
        b = oracle.pack(s, len(s))
        assert oracle.vkmer_hash(s) == h(struct.pack(">i", len(s)) + b)


def test_uuid_layout():
    # ReadHeadInfo bit layout: offset<<40 | lib<<36 | mate<<35 | readId
    assert oracle.make_uuid(0, 0, 1, 0) == 1
    assert oracle.make_uuid(1, 0, 5, 0) == (1 << 35) | 5
    assert oracle.make_uuid(1, 3, 5, 2) == (2 << 40) | (3 << 36) | (1 << 35) | 5
    neg = ((4 | (1 << 23)) << 40) | 9  # sign bit of the 24-bit offset -> bit 63
    assert oracle.make_uuid(0, 0, 9, -4) == neg - (1 << 64)
    with pytest.raises(oracle.OracleError):
        oracle.make_uuid(0, 0, 1 << 29, 0)
    with pytest.raises(oracle.OracleError):
        oracle.make_uuid(0, 0, 1, 1 << 23)


@pytest.mark.parametrize("f,s", [
    (1.0, "1.0"), (8.0, "8.0"), (38.0, "38.0"), (1.0e7, "1.0E7"), (0.001, "0.001"),
    (1.0e-4, "1.0E-4"), (3.14, "3.14"), (0.1, "0.1"), (1.0 / 3.0, "0.33333334"),
    (123456.7, "123456.7"), (12345678.0, "1.2345678E7"), (2.5, "2.5")])
def test_java_float_to_string(f, s):
    assert oracle.float_str(f) == s


def test_union_update_small_cases():
    # a singleton list is unchanged; duplicates collapse
    assert oracle.union_update([], ["ACG"]) == ["ACG"]
    assert oracle.union_update(["ACG"], ["ACG"]) == ["ACG"]
    got = oracle.union_update(["ACG", "CCG"], ["GGA"])
    assert sorted(got) == ["ACG", "CCG", "GGA"]


SEQFILES = sorted(glob.glob(os.path.join(GOLDEN, "seqfile", "*", "part-*")))


def _parse_seqfile(data):
    assert data[:4] == b"SEQ\x06"
    off = 4
    names = []
    for _ in range(2):
        n = data[off]
        names.append(data[off + 1:off + 1 + n].decode())
        off += 1 + n
    assert data[off:off + 2] == b"\x00\x00"  # not compressed
    off += 2
    assert struct.unpack(">i", data[off:off + 4])[0] == 0  # no metadata
    off += 4
    sync = data[off:off + 16]
    off += 16
    recs = bytearray()
    syncs = []
    while off < len(data):
        (n,) = struct.unpack(">i", data[off:off + 4])
        if n == -1:
            assert data[off + 4:off + 20] == sync
            syncs.append(off)
            off += 20
            continue
        recs += data[off:off + 8 + n]
        off += 8 + n
    return names, sync, bytes(recs), syncs


@pytest.mark.parametrize("path", SEQFILES, ids=[os.path.relpath(p, GOLDEN) for p in SEQFILES])
def test_seqfile_container_rebuild(path):
    data = open(path, "rb").read()
    names, sync, recs, syncs = _parse_seqfile(data)
    rebuilt = oracle.seqfile(recs, sync, names[0], names[1])
    assert rebuilt == data
    crc = os.path.join(os.path.dirname(path), "." + os.path.basename(path) + ".crc")
    if os.path.exists(crc):
        assert oracle.crc_sidecar(data) == open(crc, "rb").read()


def test_seqfile_fixtures_have_syncs():
    assert len(SEQFILES) == 3
    assert sum(len(_parse_seqfile(open(p, "rb").read())[3]) for p in SEQFILES) >= 4


def test_fastq_to_readids():
    m1 = b"@r0\nACGT \n+\nIIII\n@r1\nGGCC\n+\nIIII\n"
    m2 = b"@r0\nTTTT\n+\nIIII\n@r1\n  AAAA\n+\nIIII\n"
    assert oracle.fastq_to_readids(m1) == b"2\tACGT\n6\tGGCC\n"
    assert oracle.fastq_to_readids(m1, m2) == b"2\tACGT\tTTTT\n6\tGGCC\tAAAA\n"
