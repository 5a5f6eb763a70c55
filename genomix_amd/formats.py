"""Node files: SequenceFile<VKmer, Node> container, Node decoding, text form.

The device build emits bare record streams (i32 recLen, i32 keyLen, key,
value).  This module is the host half of the writer the reference uses,
KmerNodePairSequenceWriterFactory (KmerNodePairSequenceWriterFactory.java:
66-94) on top of Hadoop 0.20.2 SequenceFile.Writer: file header, a sync
marker before any record that starts at least 2000 bytes after the last
sync, and the local-filesystem ``.crc`` sidecar.  It also decodes Node values
(Node.readFields, Node.java:435-456) and prints the text form of
GenomixClusterManager.copyBinAndTextToLocal (GenomixClusterManager.java:371).
"""

import os
import struct
import zlib
from dataclasses import dataclass, field
from typing import List, Optional

KEY_CLASS = "edu.uci.ics.genomix.data.types.VKmer"
VALUE_CLASS = "edu.uci.ics.genomix.data.types.Node"
SYNC_INTERVAL = 2000  # SequenceFile.SYNC_INTERVAL = 100 * (4 + 16)
EDGE_TYPES = ("FF", "FR", "RF", "RR")


def _vint_text(s: str) -> bytes:
    b = s.encode()
    if len(b) > 127:
        raise ValueError("class name too long for a one-byte vint")
    return bytes([len(b)]) + b


def iter_records(stream: bytes):
    """(key bytes, value bytes) of each record of a bare record stream."""
    off = 0
    n = len(stream)
    while off < n:
        if off + 8 > n:
            raise ValueError("truncated record stream")
        rec_len, key_len = struct.unpack(">ii", stream[off:off + 8])
        if rec_len < key_len or off + 8 + rec_len > n:
            raise ValueError("bad record")
        body = stream[off + 8:off + 8 + rec_len]
        yield body[:key_len], body[key_len:]
        off += 8 + rec_len


def seqfile_bytes(stream: bytes, sync: bytes, key_class=KEY_CLASS, value_class=VALUE_CLASS) -> bytes:
    """Header + records with sync markers, as SequenceFile.Writer.append writes
    them (checkAndWriteSync before each record)."""
    if len(sync) != 16:
        raise ValueError("sync must be 16 bytes")
    out = bytearray(b"SEQ\x06")
    out += _vint_text(key_class) + _vint_text(value_class)
    out += b"\x00\x00" + struct.pack(">i", 0) + sync
    last_sync = 0
    off = 0
    n = len(stream)
    while off < n:
        rec_len = struct.unpack(">i", stream[off:off + 4])[0]
        size = 8 + rec_len
        if len(out) >= last_sync + SYNC_INTERVAL:
            out += struct.pack(">i", -1) + sync
            last_sync = len(out)
        out += stream[off:off + size]
        off += size
    return bytes(out)


def crc_sidecar(data: bytes, bytes_per_checksum: int = 512) -> bytes:
    """Hadoop ChecksumFileSystem sidecar: "crc\\0", i32 chunk size, CRC32 per chunk."""
    out = bytearray(b"crc\x00") + struct.pack(">i", bytes_per_checksum)
    for i in range(0, len(data), bytes_per_checksum):
        out += struct.pack(">I", zlib.crc32(data[i:i + bytes_per_checksum]) & 0xFFFFFFFF)
    return bytes(out)


def write_part(directory: str, part: int, stream: bytes, sync: bytes) -> str:
    """Writes part-<p> and its .crc sidecar; returns the file path."""
    os.makedirs(directory, exist_ok=True)
    data = seqfile_bytes(stream, sync)
    path = os.path.join(directory, f"part-{part}")
    with open(path, "wb") as f:
        f.write(data)
    with open(os.path.join(directory, f".part-{part}.crc"), "wb") as f:
        f.write(crc_sidecar(data))
    return path


def read_seqfile(data: bytes):
    """Returns (key class, value class, sync, bare record stream)."""
    if data[:4] != b"SEQ\x06":
        raise ValueError("not a SequenceFile v6")
    off = 4
    names = []
    for _ in range(2):
        n = data[off]
        names.append(data[off + 1:off + 1 + n].decode())
        off += 1 + n
    if data[off:off + 2] != b"\x00\x00":
        raise ValueError("compressed SequenceFiles are not supported")
    off += 2
    nmeta = struct.unpack(">i", data[off:off + 4])[0]
    if nmeta != 0:
        raise ValueError("metadata entries are not supported")
    off += 4
    sync = data[off:off + 16]
    off += 16
    recs = bytearray()
    while off < len(data):
        rec_len = struct.unpack(">i", data[off:off + 4])[0]
        if rec_len == -1:
            off += 20
            continue
        recs += data[off:off + 8 + rec_len]
        off += 8 + rec_len
    return names[0], names[1], sync, bytes(recs)


# ---------------------------------------------------------------------------
# values
# ---------------------------------------------------------------------------

SYMBOLS = "ACGT"


def kmer_string(k: int, b: bytes) -> str:
    """KmerUtil.recoverKmerFrom: letter 0 in the low bits of the last byte."""
    v = int.from_bytes(b, "big")
    return "".join(SYMBOLS[(v >> (2 * i)) & 3] for i in range(k))


def _read_vkmer(d: bytes, off: int):
    k = struct.unpack(">i", d[off:off + 4])[0]
    nb = (k + 3) // 4 if k > 0 else 0
    return kmer_string(k, d[off + 4:off + 4 + nb]), off + 4 + nb


@dataclass
class ReadHead:
    uuid: int
    seq: str
    mate: Optional[str]

    @property
    def offset(self):
        o = (self.uuid >> 40) & 0xFFFFFF
        return -(o & 0x7FFFFF) if o & 0x800000 else o

    @property
    def mate_id(self):
        return (self.uuid >> 35) & 1

    @property
    def library(self):
        return (self.uuid >> 36) & 0xF

    @property
    def read_id(self):
        return self.uuid & ((1 << 35) - 1)

    def __str__(self):
        # ReadHeadInfo.toString (ReadHeadInfo.java:236-240)
        return (f"{self.read_id}-{self.offset}_{self.mate_id}-{self.library} readSeq: {self.seq} "
                f"mateReadSeq: {self.mate if self.mate is not None else 'null'}")


@dataclass
class Node:
    edges: List[Optional[List[str]]] = field(default_factory=lambda: [None] * 4)
    unflipped: Optional[List[ReadHead]] = None
    flipped: Optional[List[ReadHead]] = None
    internal: Optional[str] = None
    coverage: Optional[float] = None


def _read_heads(d: bytes, off: int):
    if d[off] != 1:
        raise ValueError("read-head set stored by path")
    n = struct.unpack(">i", d[off + 1:off + 5])[0]
    off += 5
    heads = []
    for _ in range(n):
        fields = d[off]
        uuid = struct.unpack(">q", d[off + 1:off + 9])[0] & 0xFFFFFFFFFFFFFFFF
        seq, off = _read_vkmer(d, off + 9)
        mate = None
        if fields & 1:
            mate, off = _read_vkmer(d, off)
        heads.append(ReadHead(uuid, seq, mate))
    return heads, off


def decode_node(d: bytes) -> Node:
    """Node.readFields (Node.java:435-456)."""
    n = Node()
    flags = d[0]
    off = 1
    for t in range(4):
        if flags & (1 << t):
            cnt = struct.unpack(">i", d[off:off + 4])[0]
            off += 4
            lst = []
            for _ in range(cnt):
                s, off = _read_vkmer(d, off)
                lst.append(s)
            n.edges[t] = lst
    if flags & 0x10:
        n.unflipped, off = _read_heads(d, off)
    if flags & 0x20:
        n.flipped, off = _read_heads(d, off)
    if flags & 0x40:
        n.internal, off = _read_vkmer(d, off)
    if flags & 0x80:
        n.coverage = struct.unpack(">f", d[off:off + 4])[0]
        off += 4
    if off != len(d):
        raise ValueError("trailing bytes in Node value")
    return n


def java_float(f: float) -> str:
    """Float.toString for the values node files hold (shortest round trip)."""
    import numpy as np
    if f == 0:
        return "0.0"
    a = abs(f)
    r = np.format_float_positional(np.float32(f), unique=True, trim="-") if 1e-3 <= a < 1e7 else None
    if r is not None:
        return r if "." in r else r + ".0"
    s = np.format_float_scientific(np.float32(f), unique=True, trim="-")
    mant, ex = s.split("e")
    if "." not in mant:
        mant += ".0"
    return f"{mant}E{int(ex)}"


def node_text(key: str, n: Node) -> str:
    """key.toString() + "\\t" + Node.toString() (Node.java:525-537)."""
    parts = []
    for t in range(4):
        lst = n.edges[t]
        parts.append(f"{EDGE_TYPES[t]}:" + ("null" if lst is None else "[" + ",".join(lst) + "]"))
    u = "null" if n.unflipped is None else "[" + ",".join(map(str, n.unflipped)) + "]"
    f = "null" if n.flipped is None else "[" + ",".join(map(str, n.flipped)) + "]"
    parts.append(f"5':{u}, ~5':{f}")
    parts.append("kmer:" + (n.internal if n.internal is not None else "null"))
    parts.append("cov:" + (java_float(n.coverage) + "x" if n.coverage is not None else "null"))
    return key + "\t{" + "\t".join(parts) + "}"


def decode_records(stream: bytes):
    """[(key string, Node)] of a bare record stream."""
    out = []
    for kb, vb in iter_records(stream):
        k = struct.unpack(">i", kb[:4])[0]
        out.append((kmer_string(k, kb[4:]), decode_node(vb)))
    return out


def text_dump(stream: bytes) -> str:
    return "".join(node_text(k, n) + "\n" for k, n in decode_records(stream))
