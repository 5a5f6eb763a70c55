// CPU oracle for the Genomix graph build and unitig compression.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, on the CPU and in plain
// C++17, the algorithm of the Java reference (uci-cbcl/genomix) so that the
// GPU implementation can be checked against it.  Only tests/, the smoke
// check and the CPU-baseline leg of bench.py may load it; the product
// library never links or calls it.
//
// The restatement follows the reference's own dataflow at P partitions:
//   parse (one record per k-mer occurrence, serialised like
//   Node.marshalToByteArray) -> stable sort by KmerPointable order ->
//   preclustered group with AggregateKmerAggregateFactory (local) ->
//   hash partition -> merge -> group again (cross) -> serialised Node
//   records, then SequenceFile framing.
// Each function cites the reference file:line it follows (paths relative to
// the reference root).  The JDK-7 java.util.HashSet iteration order that
// VKmerList.unionUpdate exposes is modelled explicitly (bucket array,
// head-inserted chains, resize only on collision).
//
// Parity status of each part (what pins it) is listed in the header of
// oracle/oracle.py.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <thread>
#include <functional>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace ora {

struct OraError : std::runtime_error {
  int code;
  OraError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum { E_INVAL = -22, E_INPUT = -5 };

// ---------------------------------------------------------------------------
// Java primitives
// ---------------------------------------------------------------------------

// Marshal.hashBytes (genomix-data/.../utils/Marshal.java:65-70) and the
// identical KmerPartitionComputerFactory.hashBytes (:31-36): h = 31*h + (int)b
// over signed bytes, int32 wrap-around.
static int32_t java_hash_bytes(const uint8_t* b, size_t n) {
  uint32_t h = 1;
  for (size_t i = 0; i < n; i++) h = h * 31u + (uint32_t)(int32_t)(int8_t)b[i];
  return (int32_t)h;
}

static void put_be32(std::vector<uint8_t>& o, uint32_t v) {
  o.push_back(v >> 24); o.push_back(v >> 16); o.push_back(v >> 8); o.push_back(v);
}
static void put_be64(std::vector<uint8_t>& o, uint64_t v) {
  for (int s = 56; s >= 0; s -= 8) o.push_back((uint8_t)(v >> s));
}
static uint32_t get_be32(const uint8_t* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static uint64_t get_be64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v = (v << 8) | p[i];
  return v;
}

// GeneCode.getCodeFromSymbol (GeneCode.java:29-50): A/a=0 C/c=1 G/g=2 T/t=3,
// anything else 0.
static inline uint8_t gene_code(uint8_t ch) {
  switch (ch) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
  }
  return 0;
}
static const char kSym[4] = {'A', 'C', 'G', 'T'};

// KmerUtil.getByteNumFromK (KmerUtil.java:21-27).
static inline int bytes_for_k(int k) { return k / 4 + (k % 4 != 0 ? 1 : 0); }

// ---------------------------------------------------------------------------
// Kmer / VKmer byte packing
// ---------------------------------------------------------------------------

// Kmer.setFromStringBytes (Kmer.java:225-242); identical loop in
// VKmer.setFromStringBytes (VKmer.java:462-479).
static void pack_forward(const uint8_t* s, size_t slen, size_t start, int k, uint8_t* out) {
  int nb = bytes_for_k(k);
  uint8_t l = 0;
  int bytecount = 0, bcount = nb - 1;
  for (size_t i = start; i < start + (size_t)k && i < slen; i++) {
    uint8_t code = gene_code(s[i]);
    l |= (uint8_t)(code << bytecount);
    bytecount += 2;
    if (bytecount == 8) { out[bcount--] = l; l = 0; bytecount = 0; }
  }
  if (bcount >= 0) out[0] = l;
}

// Kmer.setReversedFromStringBytes (Kmer.java:253-272): complement codes taken
// from the last letter backwards.
static void pack_reverse(const uint8_t* s, size_t slen, size_t start, int k, uint8_t* out) {
  int nb = bytes_for_k(k);
  uint8_t l = 0;
  int bytecount = 0, bcount = nb - 1;
  for (long i = (long)start + k - 1; i >= (long)start && (size_t)i < slen; i--) {
    uint8_t code = 3 - gene_code(s[i]);
    l |= (uint8_t)(code << bytecount);
    bytecount += 2;
    if (bytecount == 8) { out[bcount--] = l; l = 0; bytecount = 0; }
  }
  if (bcount >= 0) out[0] = l;
}

// Kmer.shiftKmerWithNextCode (Kmer.java:292-303) incl. clearLeadBit (:332-336).
static void shift_with_next(uint8_t* b, int k, uint8_t c) {
  int nb = bytes_for_k(k);
  for (int i = nb - 1; i > 0; i--) {
    uint8_t in = b[i - 1] & 0x03;
    b[i] = (uint8_t)(((b[i] >> 2) & 0x3f) | (in << 6));
  }
  int pos = ((k - 1) % 4) << 1;
  b[0] = (uint8_t)(((b[0] >> 2) & 0x3f) | (uint8_t)(c << pos));
  if (k % 4 != 0) b[0] &= (uint8_t)((1 << ((k % 4) << 1)) - 1);
}

// WritableComparator.compareBytes (Hadoop 0.20.2, un-vendored): unsigned
// lexicographic compare, then length difference.  Used by
// BinaryComparable.compareTo in the canonical choice
// (ReadsKeyValueParserFactory.java:163,181).
static int compare_bytes(const uint8_t* a, int la, const uint8_t* b, int lb) {
  int n = std::min(la, lb);
  for (int i = 0; i < n; i++) {
    if (a[i] != b[i]) return (int)a[i] - (int)b[i];
  }
  return la - lb;
}

// KmerUtil.recoverKmerFrom (KmerUtil.java:34-48).
static std::string kmer_to_string(int k, const uint8_t* b, int nb) {
  std::string s;
  if (nb - 1 < 0 || k < 1) return s;
  int byteId = nb - 1;
  uint8_t cur = b[byteId];
  for (int g = 0; g < k; g++) {
    if (g % 4 == 0 && g > 0) cur = b[--byteId];
    s.push_back(kSym[(cur >> ((g % 4) * 2)) & 3]);
  }
  return s;
}

// VKmer: BE int32 letter count followed by the packed bytes
// (VKmer.java:389-391 write, :394-396 hashCode over header+bytes).
struct VKmer {
  int32_t k = 0;
  std::vector<uint8_t> b;
  bool operator==(const VKmer& o) const { return k == o.k && b == o.b; }
  void write(std::vector<uint8_t>& o) const {
    put_be32(o, (uint32_t)k);
    o.insert(o.end(), b.begin(), b.end());
  }
  int32_t hash_code() const {
    std::vector<uint8_t> t;
    write(t);
    return java_hash_bytes(t.data(), t.size());
  }
  std::string str() const { return kmer_to_string(k, b.data(), (int)b.size()); }
  static VKmer from_string(const uint8_t* s, size_t len) {  // VKmer.setAsCopy(String) :140-142
    VKmer v;
    v.k = (int32_t)len;
    v.b.assign(bytes_for_k((int)len), 0);
    if (len) pack_forward(s, len, 0, (int)len, v.b.data());
    return v;
  }
};

static size_t read_vkmer(const uint8_t* p, size_t avail, VKmer& v) {  // VKmer.setAsCopy(byte[],int) :151-156
  if (avail < 4) throw OraError(E_INPUT, "truncated VKmer");
  v.k = (int32_t)get_be32(p);
  int nb = v.k > 0 ? bytes_for_k(v.k) : 0;
  if (avail < 4 + (size_t)nb) throw OraError(E_INPUT, "truncated VKmer body");
  v.b.assign(p + 4, p + 4 + nb);
  return 4 + nb;
}

// ---------------------------------------------------------------------------
// JDK-7 java.util.HashSet<VKmer> (HashMap, hashSeed = 0, no alternative
// hashing).  VKmerList.unionUpdate (VKmerList.java:117-133) builds one with
// initial capacity |L|+|M|, adds L then M in order and rewrites the list in
// the set's iteration order.
// ---------------------------------------------------------------------------

static inline int32_t jdk7_spread(int32_t h) {  // HashMap.hash(int) in JDK 7
  uint32_t x = (uint32_t)h;
  x ^= (x >> 20) ^ (x >> 12);
  return (int32_t)(x ^ (x >> 7) ^ (x >> 4));
}

struct Jdk7HashSet {
  struct Entry { int32_t hash; const VKmer* key; };
  // chains[i] stores the chain with its HEAD at the back (push_back == insert
  // at head, as HashMap.createEntry does).
  std::vector<std::vector<Entry>> chains;
  int cap = 1, thr = 0, size = 0;

  explicit Jdk7HashSet(int n) {
    // HashMap(int): capacity = roundUpToPowerOf2(n) (n<=1 -> 1), threshold =
    // (int)(capacity * 0.75f).
    cap = 1;
    while (cap < n) cap <<= 1;
    thr = (int)((float)cap * 0.75f);
    chains.assign(cap, {});
  }
  void add(const VKmer* x) {
    int32_t h = jdk7_spread(x->hash_code());
    int bi = h & (cap - 1);
    for (auto& e : chains[bi])
      if (e.hash == h && *e.key == *x) return;  // present: HashMap.put replaces value only
    // HashMap.addEntry: resize only when size >= threshold AND the target
    // bucket is non-empty.
    if (size >= thr && !chains[bi].empty()) {
      int ncap = cap * 2;
      std::vector<std::vector<Entry>> nt(ncap);
      // HashMap.transfer: old buckets in order, each chain head->tail, every
      // entry pushed to the head of its new chain.
      for (int j = 0; j < cap; j++) {
        auto& ch = chains[j];
        for (long t = (long)ch.size() - 1; t >= 0; t--) {
          const Entry& e = ch[t];
          nt[e.hash & (ncap - 1)].push_back(e);
        }
      }
      chains.swap(nt);
      cap = ncap;
      thr = (int)((float)cap * 0.75f);
      bi = h & (cap - 1);
    }
    chains[bi].push_back({h, x});
    size++;
  }
  std::vector<VKmer> iterate() const {  // HashIterator: buckets ascending, chain head->tail
    std::vector<VKmer> out;
    for (int j = 0; j < cap; j++)
      for (long t = (long)chains[j].size() - 1; t >= 0; t--) out.push_back(*chains[j][t].key);
    return out;
  }
};

// Deliberately different HashMap generations, used only by the tests as
// negative controls (a wrong model must fail the fixture edge-order replay):
//   1 = JDK 8 (hash h ^ h>>>16, tail insertion, resize after ++size >
//       threshold with order-preserving lo/hi split),
//   2 = JDK 6 (JDK-7 hash and head insertion, but resize after every
//       size++ >= threshold regardless of collisions).
static int g_hash_model = 0;

struct OtherHashSet {
  struct Entry { int32_t hash; const VKmer* key; };
  std::vector<std::vector<Entry>> chains;  // head at FRONT here
  int cap = 1, thr = 0, size = 0, model;
  OtherHashSet(int n, int m) : model(m) {
    cap = 1;
    while (cap < n) cap <<= 1;
    thr = (int)((float)cap * 0.75f);
    chains.assign(cap, {});
  }
  int32_t spread(int32_t h) const {
    if (model == 1) return h ^ (int32_t)((uint32_t)h >> 16);
    return jdk7_spread(h);
  }
  void rehash(int ncap) {
    std::vector<std::vector<Entry>> nt(ncap);
    for (int j = 0; j < cap; j++) {
      if (model == 1) {
        for (auto& e : chains[j]) nt[e.hash & (ncap - 1)].push_back(e);  // order kept
      } else {
        for (auto& e : chains[j]) {
          auto& d = nt[e.hash & (ncap - 1)];
          d.insert(d.begin(), e);
        }
      }
    }
    chains.swap(nt);
    cap = ncap;
    thr = (int)((float)cap * 0.75f);
  }
  void add(const VKmer* x) {
    int32_t h = spread(x->hash_code());
    int bi = h & (cap - 1);
    for (auto& e : chains[bi])
      if (e.hash == h && *e.key == *x) return;
    if (model == 1) {
      chains[bi].push_back({h, x});
      if (++size > thr) rehash(cap * 2);
    } else {
      chains[bi].insert(chains[bi].begin(), {h, x});
      if (size++ >= thr) rehash(cap * 2);
    }
  }
  std::vector<VKmer> iterate() const {
    std::vector<VKmer> out;
    for (int j = 0; j < cap; j++)
      for (auto& e : chains[j]) out.push_back(*e.key);
    return out;
  }
};

// VKmerList.unionUpdate (VKmerList.java:117-133).
static std::vector<VKmer> union_update(const std::vector<VKmer>& L, const std::vector<VKmer>& M) {
  if (g_hash_model != 0) {
    OtherHashSet s((int)(L.size() + M.size()), g_hash_model);
    for (auto& x : L) s.add(&x);
    for (auto& x : M) s.add(&x);
    return s.iterate();
  }
  Jdk7HashSet s((int)(L.size() + M.size()));
  for (auto& x : L) s.add(&x);
  for (auto& x : M) s.add(&x);
  return s.iterate();
}

// ---------------------------------------------------------------------------
// ReadHeadInfo / ReadHeadSet
// ---------------------------------------------------------------------------

struct ReadHead {
  int64_t uuid = 0;
  VKmer seq;
  bool has_mate = false;  // mateReadSequence != null && length > 0
  VKmer mate;

  // ReadHeadInfo bit layout (ReadHeadInfo.java:17-34): offset<<40 | lib<<36 |
  // mate<<35 | readId.
  int offset() const {  // getOffset :168-176
    int off = (int)((uint64_t)uuid >> 40);
    if (off & (1 << 23)) off = -(off & ((1 << 23) - 1));
    return off;
  }
  int lib() const { return (int)(((uint64_t)uuid >> 36) & 0xF); }
  int mate_id() const { return (int)(((uint64_t)uuid >> 35) & 1); }
  int64_t read_id() const { return (int64_t)((uint64_t)uuid & ((1ull << 35) - 1)); }
};

// ReadHeadInfo.makeUUID (ReadHeadInfo.java:100-127), including its range
// checks.  Note the readId check masks with ~(-1L << 29): readIds must fit 29
// bits even though 35 are reserved.
static int64_t make_uuid(int mate, int lib, int64_t readId, int offset) {
  if (readId != (readId & (int64_t)~(~0ull << 29)))
    throw OraError(E_INPUT, "readId " + std::to_string(readId) + " does not fit the ReadHeadInfo uuid");
  const int maxOff = (1 << 23) - 1;
  if (std::abs(offset) > maxOff) throw OraError(E_INPUT, "read head offset out of range");
  if (offset < 0) offset = -offset | (maxOff + 1);
  return (int64_t)(((uint64_t)(uint32_t)offset << 40) + ((uint64_t)lib << 36) + ((uint64_t)mate << 35) +
                   (uint64_t)readId);
}

// ReadHeadInfo.compareTo (:247-264): offset, library, mate, readId.
static int head_cmp(const ReadHead& a, const ReadHead& b) {
  if (a.offset() != b.offset()) return a.offset() < b.offset() ? -1 : 1;
  if (a.lib() != b.lib()) return a.lib() < b.lib() ? -1 : 1;
  if (a.mate_id() != b.mate_id()) return a.mate_id() < b.mate_id() ? -1 : 1;
  if (a.read_id() != b.read_id()) return a.read_id() < b.read_id() ? -1 : 1;
  return 0;
}

// java.util.TreeSet semantics used by ExternalableTreeSet.add/union
// (ExternalableTreeSet.java:85-92,138-145): insertion keeps an existing equal
// element.
static void treeset_add(std::vector<ReadHead>& s, const ReadHead& h) {
  auto it = std::lower_bound(s.begin(), s.end(), h,
                             [](const ReadHead& a, const ReadHead& b) { return head_cmp(a, b) < 0; });
  if (it != s.end() && head_cmp(*it, h) == 0) return;
  s.insert(it, h);
}

// ReadHeadInfo.write (:205-212).
static void write_head(std::vector<uint8_t>& o, const ReadHead& h) {
  o.push_back(h.has_mate ? 1 : 0);
  put_be64(o, (uint64_t)h.uuid);
  h.seq.write(o);
  if (h.has_mate) h.mate.write(o);
}

// ExternalableTreeSet.write with writeEntireBody forced true (:236-267):
// boolean true, int size, elements in TreeSet order.
static void write_headset(std::vector<uint8_t>& o, const std::vector<ReadHead>& s) {
  o.push_back(1);
  put_be32(o, (uint32_t)s.size());
  for (auto& h : s) write_head(o, h);
}

static size_t read_headset(const uint8_t* p, size_t avail, std::vector<ReadHead>& s) {
  size_t off = 0;
  if (avail < 1) throw OraError(E_INPUT, "truncated read-head set");
  bool whole = p[off++] != 0;
  if (!whole) throw OraError(E_INPUT, "read-head set stored by path is not supported");
  if (avail < off + 4) throw OraError(E_INPUT, "truncated read-head set");
  uint32_t n = get_be32(p + off);
  off += 4;
  s.clear();
  for (uint32_t i = 0; i < n; i++) {
    ReadHead h;
    if (avail < off + 9) throw OraError(E_INPUT, "truncated read head");
    uint8_t fields = p[off++];
    h.uuid = (int64_t)get_be64(p + off);
    off += 8;
    off += read_vkmer(p + off, avail - off, h.seq);
    if (fields & 1) {
      off += read_vkmer(p + off, avail - off, h.mate);
      h.has_mate = true;  // readFields creates the mate object when present
    }
    treeset_add(s, h);
  }
  return off;
}

// ---------------------------------------------------------------------------
// Node (Node.java:123-128, write :408-427, active fields :466-487)
// ---------------------------------------------------------------------------

enum { FF = 0, FR = 1, RF = 2, RR = 3 };

struct Node {
  std::vector<VKmer> edges[4];
  bool edges_present[4] = {false, false, false, false};  // null vs non-null list
  std::vector<ReadHead> unflipped, flipped;
  bool has_unflipped = false, has_flipped = false;
  VKmer internal;  // k == 0 -> absent
  bool has_internal = false;
  bool has_cov = false;
  float cov = 0.f;

  void write(std::vector<uint8_t>& o) const {
    uint8_t f = 0;
    for (int e = 0; e < 4; e++)
      if (edges_present[e] && !edges[e].empty()) f |= (uint8_t)(1 << e);
    if (has_unflipped && !unflipped.empty()) f |= 1 << 4;
    if (has_flipped && !flipped.empty()) f |= 1 << 5;
    if (has_internal && internal.k > 0) f |= 1 << 6;
    if (has_cov) f |= 1 << 7;
    o.push_back(f);
    for (int e = 0; e < 4; e++) {
      if (f & (1 << e)) {  // VKmerList.write (VKmerList.java:314-316)
        put_be32(o, (uint32_t)edges[e].size());
        for (auto& v : edges[e]) v.write(o);
      }
    }
    if (f & (1 << 4)) write_headset(o, unflipped);
    if (f & (1 << 5)) write_headset(o, flipped);
    if (f & (1 << 6)) internal.write(o);
    if (f & (1 << 7)) {
      uint32_t u;
      std::memcpy(&u, &cov, 4);
      put_be32(o, u);
    }
  }

  // Node.readFields / setAsCopy (:336-360, :435-456).
  size_t read(const uint8_t* p, size_t avail) {
    *this = Node();
    if (avail < 1) throw OraError(E_INPUT, "truncated Node");
    size_t off = 0;
    uint8_t f = p[off++];
    for (int e = 0; e < 4; e++) {
      if (f & (1 << e)) {
        if (avail < off + 4) throw OraError(E_INPUT, "truncated edge list");
        uint32_t n = get_be32(p + off);
        off += 4;
        edges_present[e] = true;
        for (uint32_t i = 0; i < n; i++) {
          VKmer v;
          off += read_vkmer(p + off, avail - off, v);
          edges[e].push_back(v);
        }
      }
    }
    if (f & (1 << 4)) { has_unflipped = true; off += read_headset(p + off, avail - off, unflipped); }
    if (f & (1 << 5)) { has_flipped = true; off += read_headset(p + off, avail - off, flipped); }
    if (f & (1 << 6)) { has_internal = true; off += read_vkmer(p + off, avail - off, internal); }
    if (f & (1 << 7)) {
      if (avail < off + 4) throw OraError(E_INPUT, "truncated coverage");
      uint32_t u = get_be32(p + off);
      std::memcpy(&cov, &u, 4);
      has_cov = true;
      off += 4;
    }
    return off;
  }
};

// Java Float.toString for the values the text dumps hold: shortest decimal
// that reads back as the same float, plain notation in [1e-3, 1e7),
// computerised scientific notation otherwise.
static std::string java_float_str(float f) {
  if (f != f) return "NaN";
  if (f == 0.f) return std::signbit(f) ? "-0.0" : "0.0";
  if (f > 3.4e38f * 10) return "Infinity";
  char buf[64];
  int prec = 1;
  for (; prec <= 9; prec++) {
    snprintf(buf, sizeof buf, "%.*e", prec - 1, (double)f);
    if (strtof(buf, nullptr) == f) break;
  }
  // buf = d.ddde[+-]XX
  std::string m(buf);
  size_t epos = m.find('e');
  int ex = atoi(m.c_str() + epos + 1);
  std::string digits;
  bool neg = m[0] == '-';
  for (size_t i = neg ? 1 : 0; i < epos; i++)
    if (m[i] != '.') digits.push_back(m[i]);
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  std::string out = neg ? "-" : "";
  double a = std::fabs((double)f);
  if (a >= 1e-3 && a < 1e7) {
    if (ex >= 0) {
      std::string ip = digits.substr(0, std::min((size_t)ex + 1, digits.size()));
      while ((int)ip.size() < ex + 1) ip.push_back('0');
      std::string fp = (int)digits.size() > ex + 1 ? digits.substr(ex + 1) : "0";
      out += ip + "." + fp;
    } else {
      out += "0.";
      for (int i = 0; i < -ex - 1; i++) out.push_back('0');
      out += digits;
    }
  } else {
    out += digits.substr(0, 1) + "." + (digits.size() > 1 ? digits.substr(1) : "0") + "E" + std::to_string(ex);
  }
  return out;
}

// ReadHeadInfo.toString (:236-240); ReadHeadSet.toString (ReadHeadSet.java:46-58).
static std::string head_str(const ReadHead& h) {
  return std::to_string(h.read_id()) + "-" + std::to_string(h.offset()) + "_" + std::to_string(h.mate_id()) + "-" +
         std::to_string(h.lib()) + " readSeq: " + h.seq.str() + " mateReadSeq: " + (h.has_mate ? h.mate.str() : "null");
}
static std::string headset_str(const std::vector<ReadHead>& s) {
  std::string o = "[";
  for (size_t i = 0; i < s.size(); i++) {
    if (i) o += ",";
    o += head_str(s[i]);
  }
  return o + "]";
}

// Node.toString (Node.java:525-537) + VKmerList.toString (VKmerList.java:349-362).
static std::string node_str(const Node& n) {
  static const char* et[4] = {"FF", "FR", "RF", "RR"};
  std::string o = "{";
  for (int e = 0; e < 4; e++) {
    o += et[e];
    o += ":";
    if (!n.edges_present[e]) {
      o += "null";
    } else {
      o += "[";
      for (size_t i = 0; i < n.edges[e].size(); i++) {
        if (i) o += ",";
        o += n.edges[e][i].str();
      }
      o += "]";
    }
    o += "\t";
  }
  o += "5':" + (n.has_unflipped ? headset_str(n.unflipped) : std::string("null"));
  o += ", ~5':" + (n.has_flipped ? headset_str(n.flipped) : std::string("null")) + "\t";
  o += "kmer:" + (n.has_internal ? n.internal.str() : std::string("null")) + "\t";
  o += "cov:" + (n.has_cov ? java_float_str(n.cov) + "x" : std::string("null")) + "}";
  return o;
}

// ---------------------------------------------------------------------------
// Parser: ReadsKeyValueParserFactory.parse / SplitReads
// ---------------------------------------------------------------------------

struct Record {       // one Hyracks tuple: (Kmer bytes, Node bytes)
  std::vector<uint8_t> key;   // nb bytes
  std::vector<uint8_t> node;  // Node.marshalToByteArray
};

struct ParseCtx {
  int k;
  std::vector<Record>* out;
};

static bool valid_acgt(const uint8_t* s, size_t n) {  // "[ACGTacgt]+".matches
  if (n == 0) return false;
  for (size_t i = 0; i < n; i++) {
    switch (s[i]) {
      case 'A': case 'C': case 'G': case 'T': case 'a': case 'c': case 'g': case 't': break;
      default: return false;
    }
  }
  return true;
}

// Long.parseLong: optional sign then decimal digits, must fit int64.
static int64_t java_parse_long(const uint8_t* s, size_t n, bool* ok) {
  *ok = false;
  if (n == 0) return 0;
  size_t i = 0;
  bool neg = false;
  if (s[0] == '-' || s[0] == '+') { neg = s[0] == '-'; i = 1; if (n == 1) return 0; }
  unsigned __int128 v = 0;
  for (; i < n; i++) {
    if (s[i] < '0' || s[i] > '9') return 0;
    v = v * 10 + (s[i] - '0');
    if (v > ((unsigned __int128)1 << 63)) return 0;
  }
  if (!neg && v == ((unsigned __int128)1 << 63)) return 0;
  *ok = true;
  return neg ? -(int64_t)(v - 1) - 1 : (int64_t)v;
}

// SplitReads (ReadsKeyValueParserFactory.java:150-196) with
// setEdgesForCurAndNext (:209-233) and writeToFrame/InsertToFrame (:198-254).
static void split_reads(ParseCtx& c, const ReadHead& head_template, const uint8_t* r, size_t L) {
  const int k = c.k, nb = bytes_for_k(k);
  if ((size_t)k >= L)
    throw OraError(E_INPUT, "kmersize (k=" + std::to_string(k) + ") is larger than the read length (" +
                                std::to_string(L) + ")");
  std::vector<uint8_t> curF(nb), curR(nb), nxtF(nb), nxtR(nb);
  Node cur, nxt;
  cur.has_cov = true; cur.cov = 1.f;
  pack_forward(r, L, 0, k, curF.data());
  pack_reverse(r, L, 0, k, curR.data());
  bool curFwd = compare_bytes(curF.data(), nb, curR.data(), nb) <= 0;
  ReadHead h = head_template;
  if (curFwd) {
    cur.has_unflipped = true;
    treeset_add(cur.unflipped, h);
  } else {
    h.uuid = make_uuid(h.mate_id(), h.lib(), h.read_id(), k - 1);  // resetOffset(K-1)
    cur.has_flipped = true;
    treeset_add(cur.flipped, h);
  }
  nxt = Node(); nxt.has_cov = true; nxt.cov = 1.f;
  nxtF = curF;
  auto vk = [&](const std::vector<uint8_t>& b) { VKmer v; v.k = k; v.b = b; return v; };
  auto emit = [&](const std::vector<uint8_t>& F, const std::vector<uint8_t>& R, bool fwd, const Node& n) {
    Record rec;
    rec.key = fwd ? F : R;
    n.write(rec.node);
    c.out->push_back(std::move(rec));
  };
  for (size_t i = k; i < L; i++) {
    shift_with_next(nxtF.data(), k, gene_code(r[i]));
    pack_reverse(r, L, i - k + 1, k, nxtR.data());
    bool nxtFwd = compare_bytes(nxtF.data(), nb, nxtR.data(), nb) <= 0;
    if (curFwd && nxtFwd) {
      cur.edges_present[FF] = true; cur.edges[FF].push_back(vk(nxtF));
      nxt.edges_present[RR] = true; nxt.edges[RR].push_back(vk(curF));
    } else if (curFwd && !nxtFwd) {
      cur.edges_present[FR] = true; cur.edges[FR].push_back(vk(nxtR));
      nxt.edges_present[FR] = true; nxt.edges[FR].push_back(vk(curF));
    } else if (!curFwd && nxtFwd) {
      cur.edges_present[RF] = true; cur.edges[RF].push_back(vk(nxtF));
      nxt.edges_present[RF] = true; nxt.edges[RF].push_back(vk(curR));
    } else {
      cur.edges_present[RR] = true; cur.edges[RR].push_back(vk(nxtR));
      nxt.edges_present[FF] = true; nxt.edges[FF].push_back(vk(curR));
    }
    emit(curF, curR, curFwd, cur);
    curF = nxtF; curR = nxtR; cur = nxt; curFwd = nxtFwd;
    nxt = Node(); nxt.has_cov = true; nxt.cov = 1.f;
  }
  emit(curF, curR, curFwd, cur);
}

// ReadsKeyValueParserFactory.parse (:95-148).  libraryId is always 0: the
// pattern's group(0) is read without matches() and throws, which is caught.
static void parse_line(ParseCtx& c, const uint8_t* line, size_t n) {
  // String.split("\\t"): trailing empty strings removed.
  std::vector<std::pair<size_t, size_t>> f;
  size_t st = 0;
  for (size_t i = 0; i <= n; i++) {
    if (i == n || line[i] == '\t') { f.push_back({st, i - st}); st = i + 1; }
  }
  while (!f.empty() && f.back().second == 0) f.pop_back();
  if (f.empty()) f.push_back({0, 0});  // "".split gives [""]
  if (f.size() != 2 && f.size() != 3)
    throw OraError(E_INPUT, "input format is not correct! only support id'\\t'readSeq'\\t'mateReadSeq or id'\\t'readSeq");
  bool ok;
  int64_t readId = java_parse_long(line + f[0].first, f[0].second, &ok);
  if (!ok) throw OraError(E_INPUT, "bad read id");
  const uint8_t* m0 = line + f[1].first;
  size_t l0 = f[1].second;
  const uint8_t* m1 = f.size() == 3 ? line + f[2].first : nullptr;
  size_t l1 = f.size() == 3 ? f[2].second : 0;
  if (valid_acgt(m0, l0)) {
    ReadHead h;
    h.uuid = make_uuid(0, 0, readId, 0);
    h.seq = VKmer::from_string(m0, l0);
    if (m1) { h.mate = VKmer::from_string(m1, l1); h.has_mate = l1 > 0; }
    split_reads(c, h, m0, l0);
  }
  if (m1 && valid_acgt(m1, l1)) {
    ReadHead h;
    h.uuid = make_uuid(1, 0, readId, 0);
    h.seq = VKmer::from_string(m1, l1);
    h.mate = VKmer::from_string(m0, l0);
    h.has_mate = l0 > 0;
    split_reads(c, h, m1, l1);
  }
}

// Hadoop LineReader: lines end at '\n', '\r' or "\r\n"; no record after the
// final terminator.
static void for_each_line(const uint8_t* t, size_t len,
                          const std::function<void(size_t, const uint8_t*, size_t)>& fn) {
  size_t i = 0;
  while (i < len) {
    size_t j = i;
    while (j < len && t[j] != '\n' && t[j] != '\r') j++;
    fn(i, t + i, j - i);
    if (j < len && t[j] == '\r' && j + 1 < len && t[j + 1] == '\n') j++;
    i = j + 1;
  }
}

// ---------------------------------------------------------------------------
// Dataflow (JobGenBuildBrujinGraph.java:92-117)
// ---------------------------------------------------------------------------

// KmerPointable.compareTo (KmerPointable.java:94-107): length, then bytes
// from the LAST to the first, unsigned.
static int kmer_pointable_cmp(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b) {
  if (a.size() != b.size()) return (int)a.size() - (int)b.size();
  for (long i = (long)a.size() - 1; i >= 0; i--)
    if (a[i] != b[i]) return (int)a[i] - (int)b[i];
  return 0;
}

// AggregateKmerAggregateFactory init/aggregate (:93-144).
struct AggState {
  Node n;
};
static void agg_init(AggState& s, const Node& rec) {
  s.n = Node();
  for (int e = 0; e < 4; e++) {
    s.n.edges[e] = union_update({}, rec.edges[e]);
    s.n.edges_present[e] = true;
  }
  s.n.has_unflipped = s.n.has_flipped = true;
  for (auto& h : rec.unflipped) treeset_add(s.n.unflipped, h);
  for (auto& h : rec.flipped) treeset_add(s.n.flipped, h);
  s.n.has_cov = true;
  s.n.cov = rec.cov;
}
static void agg_step(AggState& s, const Node& rec) {
  for (int e = 0; e < 4; e++) s.n.edges[e] = union_update(s.n.edges[e], rec.edges[e]);
  for (auto& h : rec.unflipped) treeset_add(s.n.unflipped, h);
  for (auto& h : rec.flipped) treeset_add(s.n.flipped, h);
  s.n.cov = s.n.cov + rec.cov;
}

// ExternalSort (stable, FrameSorterMergeSort.java:153-185) +
// PreclusteredGroup (PreclusteredGroupWriter.java:86-124) with the
// aggregator.  Records within a group are visited in input order.
static std::vector<Record> sort_and_group(std::vector<Record>& in) {
  std::vector<size_t> idx(in.size());
  for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t a, size_t b) { return kmer_pointable_cmp(in[a].key, in[b].key) < 0; });
  std::vector<Record> out;
  size_t i = 0;
  Node tmp;
  while (i < idx.size()) {
    size_t j = i;
    AggState st;
    tmp.read(in[idx[i]].node.data(), in[idx[i]].node.size());
    agg_init(st, tmp);
    for (j = i + 1; j < idx.size() && kmer_pointable_cmp(in[idx[j]].key, in[idx[i]].key) == 0; j++) {
      tmp.read(in[idx[j]].node.data(), in[idx[j]].node.size());
      agg_step(st, tmp);
    }
    Record r;
    r.key = in[idx[i]].key;
    st.n.write(r.node);  // outputFinalResult (:185-214)
    out.push_back(std::move(r));
    i = j;
  }
  return out;
}

// KmerPartitionComputerFactory.partition (:41-53).
static int partition_of(const std::vector<uint8_t>& key, int P) {
  int32_t h = java_hash_bytes(key.data(), key.size());
  if (h < 0) h = -(h + 1);
  return h % P;
}

struct BuildResult {
  int k = 0;
  std::vector<std::vector<Record>> parts;  // records in output order
};

// The whole graph build.  Producers: the input lines are split into
// `nsplits` producers by the byte offsets in `split_starts` (a line belongs to
// the split holding its first byte), each producer sorts and groups locally;
// every local record goes to partition hash%P; the consumer merges producer
// streams in key order with ties taken in ascending producer order and
// groups again (cross aggregation).  local_only skips the cross step and
// returns producer outputs as "parts" (used to check fixture edge orders).
static BuildResult build(const uint8_t* text, size_t len, int k, int P, const int64_t* split_starts, int nsplits,
                         bool local_only) {
  if (k < 3) throw OraError(E_INVAL, "kmerLength must be at least 3!");  // GenomixJobConf.java:422-423
  if (k > 63) throw OraError(E_INVAL, "k > 63 is not supported");
  if (P < 1) throw OraError(E_INVAL, "nparts must be >= 1");
  if (nsplits < 1) nsplits = 1;
  std::vector<std::vector<Record>> prod(nsplits);
  for_each_line(text, len, [&](size_t start, const uint8_t* line, size_t n) {
    int p = 0;
    if (split_starts)
      while (p + 1 < nsplits && (size_t)split_starts[p + 1] <= start) p++;
    ParseCtx c{k, &prod[p]};
    parse_line(c, line, n);
  });
  std::vector<std::vector<Record>> local(nsplits);
  for (int p = 0; p < nsplits; p++) local[p] = sort_and_group(prod[p]);
  BuildResult res;
  res.k = k;
  if (local_only) {
    res.parts = std::move(local);
    return res;
  }
  res.parts.resize(P);
  for (int part = 0; part < P; part++) {
    // Gather this partition's records from each producer (each already in
    // key order) and merge them; stable_sort over the concatenation in
    // producer order gives "ties in ascending producer order".
    std::vector<Record> merged;
    for (int p = 0; p < nsplits; p++)
      for (auto& r : local[p])
        if (partition_of(r.key, P) == part) merged.push_back(r);
    res.parts[part] = sort_and_group(merged);
  }
  return res;
}

// The same dataflow with its stages on threads, as Hyracks runs it with
// threadsPerMachine = P on one node (JobGen.java:60-72: P partitions; one
// reader split per partition): P producers parse and group their byte range
// of the input concurrently (local aggregation), then P consumers merge and
// group their partition concurrently (cross aggregation), each taking the
// producers' records in ascending producer order.  The result equals
// build() with the same split starts.  Used for the CPU baseline only.
static BuildResult build_threaded(const uint8_t* text, size_t len, int k, int P, int nthreads) {
  if (k < 3) throw OraError(E_INVAL, "kmerLength must be at least 3!");
  if (k > 63) throw OraError(E_INVAL, "k > 63 is not supported");
  if (P < 1) throw OraError(E_INVAL, "nparts must be >= 1");
  if (nthreads < 1) nthreads = 1;
  // split starts: equal byte ranges, each moved to the next line start
  std::vector<size_t> start(P + 1, len);
  start[0] = 0;
  for (int p = 1; p < P; p++) {
    size_t s = std::max(start[p - 1], len * (size_t)p / (size_t)P);
    // a split starts right after a line terminator ("\r\n" counts as one)
    while (s > start[p - 1] && s < len &&
           ((text[s - 1] != '\n' && text[s - 1] != '\r') || (text[s - 1] == '\r' && text[s] == '\n')))
      s++;
    start[p] = s;
  }
  auto run = [&](int ntasks, const std::function<void(int)>& task) {
    std::atomic<int> next{0};
    std::vector<std::exception_ptr> err(ntasks);
    std::vector<std::thread> pool;
    for (int t = 0; t < std::min(nthreads, ntasks); t++)
      pool.emplace_back([&] {
        for (int i; (i = next.fetch_add(1)) < ntasks;) {
          try {
            task(i);
          } catch (...) {
            err[i] = std::current_exception();
          }
        }
      });
    for (auto& th : pool) th.join();
    for (auto& e : err)
      if (e) std::rethrow_exception(e);
  };
  std::vector<std::vector<Record>> local(P);
  run(P, [&](int p) {
    std::vector<Record> prod;
    ParseCtx c{k, &prod};
    for_each_line(text + start[p], start[p + 1] - start[p],
                  [&](size_t, const uint8_t* line, size_t n) { parse_line(c, line, n); });
    local[p] = sort_and_group(prod);
  });
  BuildResult res;
  res.k = k;
  res.parts.resize(P);
  run(P, [&](int part) {
    std::vector<Record> merged;
    for (int p = 0; p < P; p++)
      for (auto& r : local[p])
        if (partition_of(r.key, P) == part) merged.push_back(r);
    res.parts[part] = sort_and_group(merged);
  });
  return res;
}

// Record stream of one part as SequenceFile records without header or syncs:
// i32 recLen, i32 keyLen, VKmer key, Node value
// (KmerNodePairSequenceWriterFactory.java:79-94; SequenceFile.Writer.append).
static std::vector<uint8_t> part_records(const BuildResult& r, int part) {
  std::vector<uint8_t> o;
  for (auto& rec : r.parts[part]) {
    VKmer key;
    key.k = r.k;
    key.b = rec.key;
    std::vector<uint8_t> kb;
    key.write(kb);
    put_be32(o, (uint32_t)(kb.size() + rec.node.size()));
    put_be32(o, (uint32_t)kb.size());
    o.insert(o.end(), kb.begin(), kb.end());
    o.insert(o.end(), rec.node.begin(), rec.node.end());
  }
  return o;
}

// ---------------------------------------------------------------------------
// SequenceFile v6 container (Hadoop 0.20.2 SequenceFile.Writer) and the
// local-filesystem .crc sidecar (ChecksumFileSystem).
// ---------------------------------------------------------------------------

static void write_vint_text(std::vector<uint8_t>& o, const std::string& s) {
  // WritableUtils.writeVInt for 0 <= n <= 127 is one byte; class names are short.
  if (s.size() > 127) throw OraError(E_INVAL, "class name too long");
  o.push_back((uint8_t)s.size());
  o.insert(o.end(), s.begin(), s.end());
}

static std::vector<uint8_t> seqfile_wrap(const uint8_t* recs, size_t n, const uint8_t sync[16],
                                         const char* keycls, const char* valcls) {
  std::vector<uint8_t> o = {'S', 'E', 'Q', 6};
  write_vint_text(o, keycls);
  write_vint_text(o, valcls);
  o.push_back(0);  // compressed
  o.push_back(0);  // block compressed
  put_be32(o, 0);  // metadata count
  o.insert(o.end(), sync, sync + 16);
  size_t lastSync = 0, off = 0;
  while (off < n) {
    if (off + 8 > n) throw OraError(E_INPUT, "truncated record stream");
    uint32_t recLen = get_be32(recs + off);
    size_t sz = 8 + (size_t)recLen;
    if (off + sz > n) throw OraError(E_INPUT, "truncated record");
    // checkAndWriteSync: pos >= lastSyncPos + SYNC_INTERVAL (100 * 20)
    if (o.size() >= lastSync + 2000) {
      put_be32(o, 0xffffffffu);
      o.insert(o.end(), sync, sync + 16);
      lastSync = o.size();
    }
    o.insert(o.end(), recs + off, recs + off + sz);
    off += sz;
  }
  return o;
}

static uint32_t crc32_ieee(const uint8_t* p, size_t n) {
  static uint32_t tab[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int j = 0; j < 8; j++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      tab[i] = c;
    }
    init = true;
  }
  uint32_t c = 0xffffffffu;
  for (size_t i = 0; i < n; i++) c = tab[(c ^ p[i]) & 0xff] ^ (c >> 8);
  return c ^ 0xffffffffu;
}

static std::vector<uint8_t> crc_sidecar(const uint8_t* p, size_t n) {
  std::vector<uint8_t> o = {'c', 'r', 'c', 0};
  put_be32(o, 512);
  for (size_t off = 0; off < n; off += 512) put_be32(o, crc32_ieee(p + off, std::min((size_t)512, n - off)));
  return o;
}

// Text dump of a record stream: key.toString() + "\t" + value.toString()
// (GenomixClusterManager.copyBinAndTextToLocal, GenomixClusterManager.java:371).
static std::string dump_text(const uint8_t* recs, size_t n) {
  std::string o;
  size_t off = 0;
  while (off < n) {
    if (off + 8 > n) throw OraError(E_INPUT, "truncated record stream");
    uint32_t recLen = get_be32(recs + off);
    uint32_t keyLen = get_be32(recs + off + 4);
    const uint8_t* kp = recs + off + 8;
    if (off + 8 + recLen > n || keyLen > recLen) throw OraError(E_INPUT, "bad record");
    VKmer key;
    read_vkmer(kp, keyLen, key);
    Node node;
    node.read(kp + keyLen, recLen - keyLen);
    o += key.str() + "\t" + node_str(node) + "\n";
    off += 8 + recLen;
  }
  return o;
}

// GenomixDriver.convertAndUploadFastqToHDFS (GenomixDriver.java:665-714):
// readId = line number (1-based) of the sequence line = 4i+2; lines trimmed
// (String.trim removes chars <= ' ').
static std::string fastq_to_readids(const uint8_t* m1, size_t n1, const uint8_t* m2, size_t n2, bool paired) {
  auto lines = [](const uint8_t* t, size_t n) {
    std::vector<std::pair<const uint8_t*, size_t>> v;
    for_each_line(t, n, [&](size_t, const uint8_t* l, size_t ln) { v.push_back({l, ln}); });
    return v;
  };
  auto trim = [](std::pair<const uint8_t*, size_t> l) {
    const uint8_t* p = l.first;
    size_t n = l.second;
    while (n && *p <= ' ') { p++; n--; }
    while (n && p[n - 1] <= ' ') n--;
    return std::string((const char*)p, n);
  };
  auto a = lines(m1, n1);
  std::string o;
  if (paired) {
    auto b = lines(m2, n2);
    if (a.size() != b.size()) throw OraError(E_INPUT, "Fastq files didn't have the same number of lines!");
    for (size_t i = 0; i < a.size(); i++)
      if (i % 4 == 1) o += std::to_string(i + 1) + "\t" + trim(a[i]) + "\t" + trim(b[i]) + "\n";
  } else {
    for (size_t i = 0; i < a.size(); i++)
      if (i % 4 == 1) o += std::to_string(i + 1) + "\t" + trim(a[i]) + "\n";
  }
  return o;
}

// ---------------------------------------------------------------------------
// Unitig compression, sequential restatement of the fixed point of P4 path
// merging (P4ForPathMergeVertex.java:21-266, BasicPathMergeVertex.java:33-219,
// DeBruijnGraphCleanVertex.java:234-245, Node.mergeWithNodeUsingTruncatedKmer
// Node.java:568-806).  A link X -(t)- Y is merged iff X's side of t has
// degree 1, Y's side of mirror(t) has degree 1, X != Y and neither node has
// itself in an edge list (tandem repeat: restrictNeighbors restricts both
// directions, BasicPathMergeVertex.java:33-73).  The random head/tail choice
// of P4 only decides which k-mer survives and the orientation; this walker
// fixes them: the survivor is the unitig's first k-mer in KmerPointable
// order, the sequence is oriented so that it reads forward, and a cycle is
// cut at the link entering it (the cycle closes on that node, as the
// tandem-repeat stop of P4 leaves it).
// ---------------------------------------------------------------------------

static std::string rc_string(const std::string& s) {
  std::string r(s.rbegin(), s.rend());
  for (auto& c : r) c = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
  return r;
}

struct GNode {
  VKmer key;
  Node node;
};

// EDGETYPE.dir / neighborDir / mirror (EDGETYPE.java:44-188): side 0 = F, 1 = R.
static inline int et_side(int t) { return (t == FF || t == FR) ? 0 : 1; }
static inline int et_mirror(int t) { return t == FF ? RR : t == RR ? FF : t; }

static std::vector<GNode> parse_graph(const uint8_t* recs, size_t n, int k) {
  std::vector<GNode> g;
  size_t off = 0;
  while (off < n) {
    if (off + 8 > n) throw OraError(E_INPUT, "truncated record stream");
    uint32_t recLen = get_be32(recs + off), keyLen = get_be32(recs + off + 4);
    if (off + 8 + recLen > n || keyLen > recLen) throw OraError(E_INPUT, "bad record");
    GNode x;
    read_vkmer(recs + off + 8, keyLen, x.key);
    x.node.read(recs + off + 8 + keyLen, recLen - keyLen);
    if (x.key.k != k) throw OraError(E_INPUT, "key length differs from k");
    g.push_back(std::move(x));
    off += 8 + recLen;
  }
  return g;
}

static std::vector<uint8_t> graph_records(const std::vector<GNode>& g) {
  std::vector<uint8_t> out;
  for (const auto& x : g) {
    std::vector<uint8_t> kb, vb;
    x.key.write(kb);
    x.node.write(vb);
    put_be32(out, (uint32_t)(kb.size() + vb.size()));
    put_be32(out, (uint32_t)kb.size());
    out.insert(out.end(), kb.begin(), kb.end());
    out.insert(out.end(), vb.begin(), vb.end());
  }
  return out;
}

// Maximal merge paths of a node graph (the P4 fixed point), see unitigs().
struct Elem { size_t node; bool fwd; };
struct Paths {
  std::map<std::vector<uint8_t>, size_t> idx;
  std::vector<long> comp_of;
  std::vector<std::vector<Elem>> comps;
};

static Paths merge_paths(const std::vector<GNode>& g) {
  Paths P;
  const size_t N = g.size();
  auto& idx = P.idx;
  for (size_t i = 0; i < N; i++) idx[g[i].key.b] = i;
  std::vector<int> deg(2 * N, 0);
  std::vector<char> tandem(N, 0);
  std::vector<long> nbr(2 * N, -1);
  std::vector<int> nbr_side(2 * N, -1);
  for (size_t i = 0; i < N; i++) {
    for (int t = 0; t < 4; t++) {
      for (auto& z : g[i].node.edges[t]) {
        deg[2 * i + et_side(t)]++;
        if (z == g[i].key) tandem[i] = 1;
      }
    }
    for (int t = 0; t < 4; t++) {
      if (deg[2 * i + et_side(t)] == 1 && g[i].node.edges[t].size() == 1) {
        auto it = idx.find(g[i].node.edges[t][0].b);
        if (it == idx.end()) throw OraError(E_INPUT, "edge to a k-mer that is not in the graph");
        nbr[2 * i + et_side(t)] = (long)it->second;
        nbr_side[2 * i + et_side(t)] = et_side(et_mirror(t));
      }
    }
  }
  auto mergeable = [&](size_t i, int d) {
    long j = nbr[2 * i + d];
    if (deg[2 * i + d] != 1 || tandem[i] || j < 0 || (size_t)j == i) return false;
    return deg[2 * j + nbr_side[2 * i + d]] == 1 && !tandem[j];
  };
  auto& comp_of = P.comp_of;
  auto& comps = P.comps;
  comp_of.assign(N, -1);
  for (size_t i = 0; i < N; i++) {
    if (comp_of[i] >= 0) continue;
    // walk exiting F (i forward) and exiting R (i reversed)
    std::vector<Elem> fwdpart{{i, true}}, backpart;
    bool cycle = false;
    size_t cur = i;
    int exit = 0;
    while (mergeable(cur, exit)) {
      size_t j = (size_t)nbr[2 * cur + exit];
      int entered = nbr_side[2 * cur + exit];
      if (j == i) { cycle = true; break; }
      cur = j;
      exit = 1 - entered;
      fwdpart.push_back({cur, exit == 0});
    }
    if (!cycle) {
      cur = i;
      exit = 1;
      while (mergeable(cur, exit)) {
        size_t j = (size_t)nbr[2 * cur + exit];
        int entered = nbr_side[2 * cur + exit];
        cur = j;
        exit = 1 - entered;
        backpart.push_back({cur, exit == 1});  // orientation relative to i-forward
      }
    }
    std::vector<Elem> path(backpart.rbegin(), backpart.rend());
    path.insert(path.end(), fwdpart.begin(), fwdpart.end());
    size_t mn = path[0].node, mpos = 0;
    for (size_t q = 0; q < path.size(); q++)
      if (path[q].node < mn) { mn = path[q].node; mpos = q; }
    if (cycle) {
      // rotate so that the smallest node comes first, reading forward
      std::vector<Elem> rot(path.begin() + mpos, path.end());
      rot.insert(rot.end(), path.begin(), path.begin() + mpos);
      if (!rot[0].fwd) {
        std::vector<Elem> rev{rot[0]};
        for (size_t q = rot.size() - 1; q >= 1; q--) rev.push_back(rot[q]);
        rot = rev;
        for (auto& e : rot) e.fwd = !e.fwd;
      }
      path = rot;
    } else if (!path[mpos].fwd) {
      std::reverse(path.begin(), path.end());
      for (auto& e : path) e.fwd = !e.fwd;
    }
    for (auto& e : path) comp_of[e.node] = (long)comps.size();
    comps.push_back(path);
  }
  return P;
}

// Coverage of a merged path: sum of node coverages / number of k-mers
// (Node.mergeCoverage is this weighted average, computed pairwise in float
// by the reference).
static float path_coverage(const std::vector<GNode>& g, const std::vector<Elem>& path) {
  double csum = 0;
  for (auto& e : path) csum += g[e.node].node.cov;
  return (float)(csum / (double)path.size());
}

// Test hook: 0 = the restatement; 1 = FR and RF swapped in the walker's
// outer-edge types; 2 = resolve_edge_through_path without the mirror of its
// first argument; 3 = the fold's read-id merge without the offset flip.
// Planted mistakes for negative controls only.
static int g_unitig_plant = 0;

static std::vector<uint8_t> unitigs(const uint8_t* recs, size_t n, int k) {
  const std::vector<GNode> g = parse_graph(recs, n, k);
  const size_t N = g.size();
  const Paths P = merge_paths(g);
  const auto& idx = P.idx;
  const auto& comp_of = P.comp_of;
  const auto& comps = P.comps;
  // position / orientation of every node in its unitig
  std::vector<size_t> pos(N);
  std::vector<char> ori(N);
  std::vector<size_t> ukey(comps.size());
  for (size_t c = 0; c < comps.size(); c++) {
    size_t mn = comps[c][0].node;
    for (size_t q = 0; q < comps[c].size(); q++) {
      pos[comps[c][q].node] = q;
      ori[comps[c][q].node] = comps[c][q].fwd;
      mn = std::min(mn, comps[c][q].node);
    }
    ukey[c] = mn;
  }
  std::vector<size_t> order(comps.size());
  for (size_t c = 0; c < comps.size(); c++) order[c] = c;
  std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return ukey[a] < ukey[b]; });
  std::vector<uint8_t> out;
  for (size_t c : order) {
    const auto& path = comps[c];
    const size_t m = path.size();
    Node u;
    // sequence: overlap-(k-1) concatenation of the oriented k-mers
    // (VKmer.mergeWithFFKmer etc., VKmer.java:525-692)
    std::string seq;
    for (size_t q = 0; q < m; q++) {
      std::string s = g[path[q].node].key.str();
      if (!path[q].fwd) s = rc_string(s);
      if (q == 0) seq = s;
      else seq.push_back(s.back());
    }
    if (m > 1) {
      u.has_internal = true;
      u.internal = VKmer::from_string((const uint8_t*)seq.data(), seq.size());
    }
    // coverage: sum of node coverages / number of k-mers (Node.mergeCoverage
    // is this weighted average, computed pairwise in float by the reference)
    u.has_cov = true;
    u.cov = path_coverage(g, path);
    // read heads: a head's offset is a letter of its node's k-mer (0 for
    // the build's unflipped heads, k - 1 for its flipped ones,
    // ReadsKeyValueParserFactory resetOffset(K-1)); in the unitig it is that
    // letter's position, and a reversed node swaps the two sets
    // (Node.mergeUnflippedAndFlippedReadIDs, Node.java:740-806)
    for (size_t q = 0; q < m; q++) {
      const Node& x = g[path[q].node].node;
      for (int f = 0; f < 2; f++) {
        const auto& set = f ? x.flipped : x.unflipped;
        for (auto h : set) {
          bool nf = path[q].fwd ? f : !f;
          int o = (int)q + (path[q].fwd ? h.offset() : k - 1 - h.offset());
          h.uuid = make_uuid(h.mate_id(), h.lib(), h.read_id(), o);
          if (nf) { u.has_flipped = true; treeset_add(u.flipped, h); }
          else { u.has_unflipped = true; treeset_add(u.unflipped, h); }
        }
      }
    }
    // outer edges of the two ends, re-targeted to neighbour unitigs
    std::set<std::pair<int, std::vector<uint8_t>>> edges;
    for (int end = 0; end < 2; end++) {
      const Elem& e = end == 0 ? path[0] : path[m - 1];
      int uletter = end == 0 ? 1 : 0;                 // 5' end -> R, 3' end -> F
      int side = (end == 0) == e.fwd ? 1 : 0;          // node side facing out
      const Node& x = g[e.node].node;
      for (int t = 0; t < 4; t++) {
        if (et_side(t) != side) continue;
        for (auto& z : x.edges[t]) {
          size_t zi = idx.at(z.b);
          bool z5 = (t == FF || t == RF);  // touches the neighbour's 5' end
          int wletter = (z5 == (bool)ori[zi]) ? 0 : 1;
          int ut = uletter == 0 ? (wletter == 0 ? FF : FR) : (wletter == 0 ? RF : RR);
          if (g_unitig_plant == 1 && (ut == FR || ut == RF)) ut = ut == FR ? RF : FR;  // planted mistake
          edges.insert({ut, g[ukey[comp_of[zi]]].key.b});
        }
      }
    }
    for (auto& pe : edges) {
      VKmer v;
      v.k = k;
      v.b = pe.second;
      u.edges_present[pe.first] = true;
      u.edges[pe.first].push_back(v);
    }
    VKmer key = g[ukey[c]].key;
    std::vector<uint8_t> kb, vb;
    key.write(kb);
    u.write(vb);
    put_be32(out, (uint32_t)(kb.size() + vb.size()));
    put_be32(out, (uint32_t)kb.size());
    out.insert(out.end(), kb.begin(), kb.end());
    out.insert(out.end(), vb.begin(), vb.end());
  }
  return out;
}

// ---------------------------------------------------------------------------
// Second restatement of unitig compression: P4's pairwise merges replayed
// one at a time.  The walker above writes each unitig's sequence, read
// heads and outer edges in closed form.  This fold instead absorbs the
// nodes of every path, one by one, into the path's survivor with the
// reference's own merge operations, each restated from its source and
// pinned by the reference's unit-test vectors:
//   EDGETYPE.resolveEdgeThroughPath (EDGETYPE.java:130-188; vectors
//     NodeFixedTest.java:128-145),
//   Node.mergeEdges (Node.java:632-707; tables NodeRandomTest.java:274-340),
//   VKmer.mergeWith{FF,FR,RF,RR}Kmer (VKmer.java:525-692; vectors
//     VKmerFixedTest.java:131-452),
//   Node.mergeUnflippedAndFlippedReadIDs (Node.java:727-806) with
//     ReadHeadSet.prependOffsets / flipOffset (ReadHeadSet.java:109-130),
//   and the neighbour update of BasicPathMergeVertex.updateNeighbors /
//     receiveUpdates (BasicPathMergeVertex.java:104-160): when B merges into
//     C, every neighbour A on B's far side replaces its edge to B by an edge
//     to C of type resolveEdgeThroughPath(B->A, B->C).
// The type of every link is read from the evolving edge lists, never
// derived from orientations, so the fold checks the walker's closed form.
// Path membership and the survivor rule are the walker's (merge_paths).
// Coverage is the walker's f32(sum / k-mers): P4 averages pairwise in float
// in a random merge order, which changes rounding only.
// ---------------------------------------------------------------------------

static inline int et_first(int t) { return t >> 1; }  // 0 = F, 1 = R
static inline int et_second(int t) { return t & 1; }
static inline int et_make(int a, int b) { return (a << 1) | b; }

// EDGETYPE.resolveEdgeThroughPath: A <-(BtoA)- B -(BtoC)-> C gives A -> C.
// The path is valid only when the inner letters of A->B and B->C agree; the
// result takes the first letter of A->B and the second of B->C.
static int resolve_edge_through_path(int BtoA, int BtoC) {
  const int AtoB = g_unitig_plant == 2 ? BtoA : et_mirror(BtoA);
  if (et_second(AtoB) != et_first(BtoC)) throw OraError(E_INVAL, "invalid link type through a path");
  return et_make(et_first(AtoB), et_second(BtoC));
}

// VKmer.mergeWith{FF,FR,RF,RR}Kmer(initialKmerSize, other) on letters: the
// two share initialKmerSize - 1 letters.  FF appends other, FR appends its
// reverse complement, RR prepends other, RF prepends its reverse complement.
static std::string vkmer_merge(int t, int initial, const std::string& self, const std::string& other) {
  const size_t ov = (size_t)(initial - 1);
  if (initial < 1 || self.size() < ov || other.size() < ov)
    throw OraError(E_INVAL, "not enough letters in the kmers to perform a merge");
  switch (t) {
    case FF: return self + other.substr(ov);
    case FR: return self + rc_string(other).substr(ov);
    case RF: {
      const std::string r = rc_string(other);
      return r.substr(0, r.size() - ov) + self;
    }
    default: return other.substr(0, other.size() - ov) + self;
  }
}

// Node.mergeEdges(edgeType, other): the side of `self` named by the first
// letter of edgeType takes over other's lists on other's far side, each
// list re-typed through the path (FF keeps FF/FR, FR maps RF->FF and
// RR->FR, RF maps FF->RF and FR->RR, RR keeps RF/RR).
static void node_merge_edges(int t, Node& self, const Node& other) {
  const int myside = et_first(t);
  const int far = 1 - et_first(et_mirror(t));  // other's side away from self
  for (int u = 0; u < 4; u++) {
    if (et_first(u) != myside) continue;
    self.edges[u].clear();
    self.edges_present[u] = false;
  }
  for (int u = 0; u < 4; u++) {
    if (et_first(u) != far || !other.edges_present[u]) continue;
    const int mine = resolve_edge_through_path(et_mirror(t), u);
    self.edges[mine] = other.edges[u];
    self.edges_present[mine] = true;
  }
}

// Node.mergeUnflippedAndFlippedReadIDs(edgeType, other, otherLength) with
// thisLength = the letters of self before the merge.
static void node_merge_heads(int t, Node& self, const Node& other, int self_len, int other_len, int K) {
  auto shifted = [](std::vector<ReadHead> s, int by, bool flip) {  // prependOffsets / flipOffset
    for (auto& h : s) {
      const int o = flip ? by - h.offset() : by + h.offset();
      h.uuid = make_uuid(h.mate_id(), h.lib(), h.read_id(), o);
    }
    return s;
  };
  auto add_all = [](Node& n, bool flipped_set, const std::vector<ReadHead>& s) {
    if (s.empty()) return;
    auto& dst = flipped_set ? n.flipped : n.unflipped;
    (flipped_set ? n.has_flipped : n.has_unflipped) = true;
    for (auto& h : s) treeset_add(dst, h);
  };
  const bool keep = g_unitig_plant == 3;  // planted: no flip of the other's offsets
  switch (t) {
    case FF:
      add_all(self, false, shifted(other.unflipped, self_len - K + 1, false));
      add_all(self, true, shifted(other.flipped, self_len - K + 1, false));
      break;
    case FR:
      add_all(self, !keep, shifted(other.unflipped, self_len - K + other_len, !keep));
      add_all(self, keep, shifted(other.flipped, self_len - K + other_len, !keep));
      break;
    case RF:
    case RR: {
      self.unflipped = shifted(self.unflipped, other_len - K + 1, false);
      self.flipped = shifted(self.flipped, other_len - K + 1, false);
      if (t == RF) {
        add_all(self, !keep, shifted(other.unflipped, other_len - 1, !keep));
        add_all(self, keep, shifted(other.flipped, other_len - 1, !keep));
      } else {
        add_all(self, false, other.unflipped);
        add_all(self, true, other.flipped);
      }
      break;
    }
  }
}

static bool list_remove(std::vector<VKmer>& l, const VKmer& x) {
  for (size_t i = 0; i < l.size(); i++)
    if (l[i] == x) { l.erase(l.begin() + (long)i); return true; }
  return false;
}
static void list_add(Node& n, int t, const VKmer& x) {  // VKmerList.unionUpdate as a set
  n.edges_present[t] = true;
  for (auto& y : n.edges[t])
    if (y == x) return;
  n.edges[t].push_back(x);
}

static std::vector<uint8_t> unitigs_p4(const uint8_t* recs, size_t n, int k) {
  const std::vector<GNode> g = parse_graph(recs, n, k);
  const size_t N = g.size();
  const Paths P = merge_paths(g);
  std::vector<Node> cur(N);
  std::vector<std::string> seq(N);
  std::vector<double> covsum(N);
  for (size_t i = 0; i < N; i++) {
    cur[i] = g[i].node;
    seq[i] = g[i].key.str();  // NodeToGenericVertexInputFormat.java:56: internal kmer := id
    covsum[i] = g[i].node.cov;
  }
  auto find_link = [&](size_t s, int side, const VKmer& b) {
    int found = -1;
    for (int t = 0; t < 4; t++)
      if (et_first(t) == side)
        for (auto& z : cur[s].edges[t])
          if (z == b) {
            if (found >= 0) throw OraError(E_INPUT, "a merge link is listed twice");
            found = t;
          }
    if (found < 0) throw OraError(E_INPUT, "a merge link is missing from the survivor's edges");
    return found;
  };
  // B merges into the survivor S through S's link of type t to B.
  auto absorb = [&](size_t s, size_t b, int t) {
    const VKmer& sk = g[s].key;
    const VKmer& bk = g[b].key;
    // B's own type towards S (the message's mergeEdgetype.mirror()).
    int btos = -1;
    for (int u = 0; u < 4; u++)
      for (auto& z : cur[b].edges[u])
        if (z == sk && et_first(u) == et_first(et_mirror(t))) btos = u;
    if (btos < 0) throw OraError(E_INPUT, "merge link not mirrored");
    // updateNeighbors / receiveUpdates: B's far-side neighbours now point at S.
    const int far = 1 - et_first(btos);
    for (int upd = 0; upd < 4; upd++) {
      if (et_first(upd) != far) continue;
      for (const VKmer& a : std::vector<VKmer>(cur[b].edges[upd])) {
        const size_t ai = P.idx.at(a.b);
        list_remove(cur[ai].edges[et_mirror(upd)], bk);
        list_add(cur[ai], resolve_edge_through_path(upd, btos), sk);
      }
    }
    // Node.mergeWithNode(t, B): edges, read ids, coverage, internal kmer.
    node_merge_edges(t, cur[s], cur[b]);
    node_merge_heads(t, cur[s], cur[b], (int)seq[s].size(), (int)seq[b].size(), k);
    covsum[s] += covsum[b];
    seq[s] = vkmer_merge(t, k, seq[s], seq[b]);
  };
  std::vector<char> survivor(N, 0);
  for (const auto& path : P.comps) {
    size_t s = path[0].node, mpos = 0;
    for (size_t q = 0; q < path.size(); q++)
      if (path[q].node < s) { s = path[q].node; mpos = q; }
    survivor[s] = 1;
    if (!path[mpos].fwd) throw OraError(E_INPUT, "survivor does not read forward");
    // extend the 3' end (F side), then the 5' end (R side)
    for (size_t q = mpos + 1; q < path.size(); q++) {
      const size_t b = path[q].node;
      absorb(s, b, find_link(s, 0, g[b].key));
    }
    for (size_t q = mpos; q-- > 0;) {
      const size_t b = path[q].node;
      absorb(s, b, find_link(s, 1, g[b].key));
    }
  }
  std::vector<uint8_t> out;
  for (size_t s = 0; s < N; s++) {
    if (!survivor[s]) continue;
    Node u = cur[s];
    const size_t m = seq[s].size() - (size_t)k + 1;
    // GenericVertexToNodeOutputFormat.java:44-47: no internal kmer at length k
    u.has_internal = m > 1;
    u.internal = m > 1 ? VKmer::from_string((const uint8_t*)seq[s].data(), seq[s].size()) : VKmer();
    u.has_cov = true;
    u.cov = (float)(covsum[s] / (double)m);
    for (int t = 0; t < 4; t++) {  // lists ascending, as the walker writes them
      std::sort(u.edges[t].begin(), u.edges[t].end(),
                [](const VKmer& a, const VKmer& b) { return a.b < b.b; });
      u.edges_present[t] = !u.edges[t].empty();
    }
    u.has_unflipped = !u.unflipped.empty();
    u.has_flipped = !u.flipped.empty();
    std::vector<uint8_t> kb, vb;
    g[s].key.write(kb);
    u.write(vb);
    put_be32(out, (uint32_t)(kb.size() + vb.size()));
    put_be32(out, (uint32_t)kb.size());
    out.insert(out.end(), kb.begin(), kb.end());
    out.insert(out.end(), vb.begin(), vb.end());
  }
  return out;
}

// Graph cleaning, restated on the node graph: the build's nodes are merged
// (merge_paths, serialised as unitigs() does), every node of a merged
// vertex the rule removes is dropped, surviving nodes lose their edges to
// dropped ones (list order kept, as VKmerList.remove), and the surviving
// graph is merged again -- the same paths, sequences, coverages, heads and
// outer edges as merging the surviving merged vertices.
static std::vector<uint8_t> clean_nodes(const uint8_t* recs, size_t n, int k,
                                        const std::function<bool(const GNode&)>& removed) {
  std::vector<GNode> g = parse_graph(recs, n, k);
  const Paths P = merge_paths(g);
  const std::vector<uint8_t> merged = unitigs(recs, n, k);
  std::set<std::vector<uint8_t>> gone;  // keys of removed merged vertices
  for (const GNode& v : parse_graph(merged.data(), merged.size(), k))
    if (removed(v)) gone.insert(v.key.b);
  std::vector<char> dead(g.size(), 0);
  for (const auto& path : P.comps) {
    size_t mn = path[0].node;
    for (auto& e : path) mn = std::min(mn, e.node);
    if (gone.count(g[mn].key.b))  // a merged vertex is keyed by its smallest k-mer
      for (auto& e : path) dead[e.node] = 1;
  }
  std::vector<GNode> keep;
  for (size_t i = 0; i < g.size(); i++) {
    if (dead[i]) continue;
    GNode x = g[i];
    for (int t = 0; t < 4; t++) {
      std::vector<VKmer> l;
      for (auto& z : x.node.edges[t])
        if (!dead[P.idx.at(z.b)]) l.push_back(z);
      x.node.edges[t] = l;
    }
    keep.push_back(std::move(x));
  }
  return graph_records(keep);  // the surviving node graph
}

static std::vector<uint8_t> clean_graph(const uint8_t* recs, size_t n, int k,
                                        const std::function<bool(const GNode&)>& removed) {
  const std::vector<uint8_t> recs2 = clean_nodes(recs, n, k, removed);
  return unitigs(recs2.data(), recs2.size(), k);
}

// REMOVE_BAD_COVERAGE then MERGE: a merged vertex whose average coverage
// is below minCoverage or above maxCoverage is removed and its neighbours
// drop the edges to it (RemoveBadCoverageVertex.java:38-70; defaults
// GenomixJobConf.java:498-502: 3.0 / +inf).
static bool bad_coverage(const GNode& v, float minc, float maxc) { return v.node.cov < minc || v.node.cov > maxc; }

static std::vector<uint8_t> clean_bad_coverage(const uint8_t* recs, size_t n, int k, float minc, float maxc) {
  return clean_graph(recs, n, k, [&](const GNode& v) { return bad_coverage(v, minc, maxc); });
}

// TIP_SINGLE_NODE then MERGE: a merged vertex with degree 1 on one side and
// 0 on the other whose sequence is at most maxLength letters long is
// removed and its neighbour drops the edge to it
// (SingleNodeTipRemoveVertex.java:47-85; Node.degree :820-828,
// getKmerLength :191-193; default maxLength = k, GenomixJobConf.java:504-505).
static bool short_tip(const GNode& v, int k, int maxlen) {
  int deg[2] = {0, 0};
  for (int t = 0; t < 4; t++) deg[et_side(t)] += (int)v.node.edges[t].size();
  const int len = v.node.has_internal ? v.node.internal.k : k;
  const bool tip = (deg[0] == 1 && deg[1] == 0) || (deg[1] == 1 && deg[0] == 0);
  return tip && len <= maxlen;
}

static std::vector<uint8_t> clean_tips(const uint8_t* recs, size_t n, int k, int maxlen) {
  // GenomixJobConf.java:449-450
  if (maxlen < k) throw OraError(E_INVAL, "tipRemove_maxLength must be at least as long as kmerLength");
  return clean_graph(recs, n, k, [&](const GNode& v) { return short_tip(v, k, maxlen); });
}

// The reference pipeline's cleaning prefix (GenomixJobConf.java:511-517):
// REMOVE_BAD_COVERAGE, MERGE, TIP_SINGLE_NODE, MERGE.
static std::vector<uint8_t> clean_coverage_then_tips(const uint8_t* recs, size_t n, int k, float minc, float maxc,
                                                     int maxlen) {
  const std::vector<uint8_t> g1 = clean_nodes(recs, n, k, [&](const GNode& v) { return bad_coverage(v, minc, maxc); });
  return clean_tips(g1.data(), g1.size(), k, maxlen);
}

// VKmer.editDistance (VKmer.java:711-731): Levenshtein distance of two
// letter strings.
static int edit_distance(const std::string& a, const std::string& b) {
  std::vector<int> prev(b.size() + 1), cur(b.size() + 1);
  for (size_t c = 0; c <= b.size(); c++) prev[c] = (int)c;
  for (size_t r = 1; r <= a.size(); r++) {
    cur[0] = (int)r;
    for (size_t c = 1; c <= b.size(); c++) {
      const int match = a[r - 1] == b[c - 1] ? 0 : 1;
      cur[c] = std::min(std::min(prev[c] + 1, cur[c - 1] + 1), prev[c - 1] + match);
    }
    std::swap(prev, cur);
  }
  return prev[b.size()];
}

// EDGETYPE.causesFlip (EDGETYPE.java:190-204).
static inline bool et_flips(int t) { return t == FR || t == RF; }

// BUBBLE then MERGE, restated on the node graph.  SimpleBubbleMergeVertex
// (SimpleBubbleMergeVertex.java:51-320) on the merged graph: a merged
// vertex with one edge on each side, to two different vertices, and a
// sequence shorter than maxLength letters is a bubble; the neighbour with
// the larger key is its major vertex, the other its minor.  Bubbles with the
// same minor and major form a set, sorted by decreasing average coverage
// (SimpleBubbleMergeMessage.SortByCoverage; Collections.sort is stable, and
// the set's order before it -- the minor's message arrival order -- is
// taken here as the merged records' order).  processSimilarSet: the first
// bubble is the top; every later one attached to the same sides of major
// and minor (isValidMajorAndMinor) whose fraction of dissimilarity to it
// (edit distance, against the reverse complement when the two are stored
// in opposite orientations, over the shorter length, VKmer.fracDissimilar
// :915-928) is at most maxDissimilarity is removed and added to the top
// (Node.addFromNode :582-614: coverage as counts, read heads at offsets
// scaled by the length ratio, ReadHeadSet.unionUpdate :80-101); the top
// leaves the set and the rest repeats.  On the node graph: the nodes of
// removed bubbles are dropped (their neighbours lose the edges to them),
// the top's first node takes their k-mer counts, and each scaled read head
// goes to the top's node covering its offset, as that node's letter.
static std::vector<uint8_t> bubble_nodes(const uint8_t* recs, size_t n, int k, float maxd, int maxlen) {
  std::vector<GNode> g = parse_graph(recs, n, k);
  const Paths P = merge_paths(g);
  const std::vector<uint8_t> merged = unitigs(recs, n, k);
  const std::vector<GNode> mv = parse_graph(merged.data(), merged.size(), k);
  std::map<std::vector<uint8_t>, size_t> comp_by_key;  // merged vertex key -> its path
  for (size_t c = 0; c < P.comps.size(); c++) {
    size_t mn = P.comps[c][0].node;
    for (auto& e : P.comps[c]) mn = std::min(mn, e.node);
    comp_by_key[g[mn].key.b] = c;
  }
  auto letters = [&](const GNode& v) { return v.node.has_internal ? v.node.internal.k : k; };
  auto seq_of = [&](const GNode& v) { return v.node.has_internal ? v.node.internal.str() : v.key.str(); };
  struct Msg {
    size_t v;      // merged vertex
    int mtb, ntb;  // major / minor to bubble edge types
    float cov;
  };
  std::map<std::pair<std::vector<uint8_t>, std::vector<uint8_t>>, std::vector<Msg>> sets;  // (minor, major)
  for (size_t v = 0; v < mv.size(); v++) {  // detectBubble :245-252 (mv in record order)
    const Node& x = mv[v].node;
    int deg[2] = {0, 0};
    for (int t = 0; t < 4; t++) deg[et_side(t)] += (int)x.edges[t].size();
    if (deg[0] != 1 || deg[1] != 1 || letters(mv[v]) >= maxlen) continue;
    int tf = x.edges[FF].empty() ? FR : FF, tr = x.edges[RF].empty() ? RR : RF;  // getSingleNeighbor :213-223
    const VKmer& kf = x.edges[tf][0];
    const VKmer& kr = x.edges[tr][0];
    if (kf == kr) continue;  // :58-60
    const bool fwd_major = compare_bytes(kf.b.data(), (int)kf.b.size(), kr.b.data(), (int)kr.b.size()) >= 0;
    const VKmer& major = fwd_major ? kf : kr;
    const VKmer& minor = fwd_major ? kr : kf;
    const int mtb = et_mirror(fwd_major ? tf : tr), ntb = et_mirror(fwd_major ? tr : tf);
    sets[{minor.b, major.b}].push_back({v, mtb, ntb, x.cov});
  }
  std::vector<char> dead(g.size(), 0);
  std::vector<float> add_cov(g.size(), 0.f);
  struct NewHead { size_t node; bool flipped; ReadHead h; };
  std::vector<NewHead> heads;
  for (auto& kv : sets) {
    std::vector<Msg> list = kv.second;
    if (list.size() < 2) continue;
    std::stable_sort(list.begin(), list.end(), [](const Msg& a, const Msg& b) { return a.cov > b.cov; });
    while (!list.empty()) {  // processSimilarSet :163-240
      const Msg top = list.front();
      list.erase(list.begin());
      const GNode& tv = mv[top.v];
      const auto& tpath = P.comps[comp_by_key.at(tv.key.b)];
      const int Lt = letters(tv);
      for (size_t j = 0; j < list.size();) {
        const Msg cur = list[j];
        if (et_side(top.mtb) != et_side(cur.mtb) || et_side(top.ntb) != et_side(cur.ntb)) {
          j++;
          continue;
        }
        const bool same = et_flips(top.mtb) == et_flips(cur.mtb);
        const GNode& cv = mv[cur.v];
        std::string cs = seq_of(cv);
        if (!same) cs = rc_string(cs);
        const int Lo = letters(cv);
        const float frac = (float)edit_distance(seq_of(tv), cs) / (float)std::min(Lt, Lo);
        if (!(frac <= maxd)) {
          j++;
          continue;
        }
        // Node.addEdges: the bubble's edges (to the same major and minor)
        // must already be the top's, as a node graph cannot hold others
        for (int t = 0; t < 4; t++) {
          const int into = same ? t : (t == FF ? RF : t == FR ? RR : t == RF ? FF : FR);
          for (auto& z : cv.node.edges[t])
            if (std::find(tv.node.edges[into].begin(), tv.node.edges[into].end(), z) == tv.node.edges[into].end())
              throw OraError(E_INPUT, "bubble merge would add an edge the top vertex lacks");
        }
        const auto& cpath = P.comps[comp_by_key.at(cv.key.b)];
        for (auto& e : cpath) {
          dead[e.node] = 1;
          add_cov[tpath[0].node] += g[e.node].node.cov;
        }
        // ReadHeadSet.unionUpdate(other, lengthFactor, flip, otherLength)
        const float factor = (float)Lt / (float)Lo;
        for (int f = 0; f < 2; f++) {
          for (const ReadHead& h : f ? cv.node.flipped : cv.node.unflipped) {
            const int o = same ? h.offset() : Lo - 1 - h.offset();
            float t = (float)(o + 1) * factor;
            t = t - factor;
            const int a = (int)t;  // letter of the top's sequence
            const bool nf = same ? f : !f;
            const size_t q = std::min((size_t)a, tpath.size() - 1);
            const bool fwd = tpath[q].fwd;
            const int rel = a - (int)q;
            NewHead nh{tpath[q].node, fwd ? nf : !nf, h};
            nh.h.uuid = make_uuid(h.mate_id(), h.lib(), h.read_id(), fwd ? rel : k - 1 - rel);
            heads.push_back(nh);
          }
        }
        list.erase(list.begin() + j);
      }
    }
  }
  for (size_t i = 0; i < g.size(); i++) {
    g[i].node.cov += add_cov[i];
  }
  for (auto& nh : heads) {
    Node& x = g[nh.node].node;
    if (nh.flipped) { x.has_flipped = true; treeset_add(x.flipped, nh.h); }
    else { x.has_unflipped = true; treeset_add(x.unflipped, nh.h); }
  }
  std::vector<GNode> keep;
  for (size_t i = 0; i < g.size(); i++) {
    if (dead[i]) continue;
    GNode x = g[i];
    for (int t = 0; t < 4; t++) {
      std::vector<VKmer> l;
      for (auto& z : x.node.edges[t])
        if (!dead[P.idx.at(z.b)]) l.push_back(z);
      x.node.edges[t] = l;
    }
    keep.push_back(std::move(x));
  }
  return graph_records(keep);
}

// The reference pipeline's cleaning (GenomixJobConf.java:511-517):
// REMOVE_BAD_COVERAGE, MERGE, TIP_SINGLE_NODE, MERGE, BUBBLE, MERGE.
static std::vector<uint8_t> bubble_nodes(const uint8_t* recs, size_t n, int k, float maxd, int maxlen);
static std::vector<uint8_t> clean_pipeline(const uint8_t* recs, size_t n, int k, float minc, float maxc, int tiplen,
                                           float maxd, int maxlen) {
  const std::vector<uint8_t> g1 = clean_nodes(recs, n, k, [&](const GNode& v) { return bad_coverage(v, minc, maxc); });
  const std::vector<uint8_t> g2 = clean_nodes(g1.data(), g1.size(), k, [&](const GNode& v) { return short_tip(v, k, tiplen); });
  const std::vector<uint8_t> g3 = bubble_nodes(g2.data(), g2.size(), k, maxd, maxlen);
  return unitigs(g3.data(), g3.size(), k);
}

static std::vector<uint8_t> bubble_merge(const uint8_t* recs, size_t n, int k, float maxd, int maxlen) {
  const std::vector<uint8_t> g1 = bubble_nodes(recs, n, k, maxd, maxlen);
  return unitigs(g1.data(), g1.size(), k);
}

// GraphStatistics.map (GraphStatistics.java:79-146) over a record stream:
// counters as JSON {"totals": {...}, "maximum": {...}, "<name>-bins":
// {"value": count}}.  updateStats (:131-139): bin count, total, and a
// maximum counter that starts at 0.
static std::string graph_stats(const uint8_t* recs, size_t n, int k) {
  const std::vector<GNode> g = parse_graph(recs, n, k);
  std::map<std::string, std::map<std::string, long long>> groups;
  auto update = [&](const std::string& name, long long v) {
    groups[name + "-bins"][std::to_string(v)] += 1;
    groups["totals"][name] += v;
    long long& mx = groups["maximum"][name];
    if (mx < v) mx = v;
  };
  static const char* kEt[4] = {"FF", "FR", "RF", "RR"};
  for (const GNode& x : g) {
    const Node& v = x.node;
    groups["totals"]["nodes"] += 1;
    const long long out = (long long)(v.edges[FF].size() + v.edges[FR].size());
    const long long in = (long long)(v.edges[RF].size() + v.edges[RR].size());
    update("degree", in + out);
    const int len = (v.has_internal && v.internal.k > 0) ? v.internal.k : x.key.k;
    update("kmerLength", len);
    const long long covr = (long long)std::floor(v.cov + 0.5f);  // Math.round(float)
    update("coverage", covr);
    update("unflippedReadIds", (long long)v.unflipped.size());
    update("flippedReadIds", (long long)v.flipped.size());
    const float cov = v.cov;
    const bool win = cov >= 0.f - 1000.f && cov <= 0.f + 1000.f;
    const long long seed = (long long)(int32_t)((uint32_t)len * (uint32_t)(v.unflipped.size() + v.flipped.size()));
    if (win) update("scaffoldSeedScore", seed);
    for (int t = 0; t < 4; t++)
      for (auto& z : v.edges[t])
        if (z == x.key) groups["totals"][std::string("selfEdge-") + kEt[t]] += 1;
    if (in == 1 && out == 1) groups["totals"]["pathNode"] += 1;
    const long long deg[2] = {out, in};
    static const char* kDir[2] = {"FORWARD", "REVERSE"};
    for (int d = 0; d < 2; d++) {
      if (deg[d] == 0) {
        groups["totals"][std::string("tips-") + kDir[d]] += 1;
      } else {
        update(std::string("kmerLength-with-") + kDir[d], len);
        update(std::string("coverage-with-") + kDir[d], covr);
        if (win) update(std::string("scaffoldSeedScore-with-") + kDir[d], seed);
      }
    }
    if (in == 0 && out == 0) groups["totals"]["tips-BOTH"] += 1;
    if ((in == 0 && out != 0) || (in != 0 && out == 0)) groups["totals"]["tips-ONE"] += 1;
  }
  std::string js = "{";
  bool fg = true;
  for (auto& gp : groups) {
    js += (fg ? "" : ",") + std::string("\"") + gp.first + "\":{";
    fg = false;
    bool fc = true;
    for (auto& c : gp.second) {
      js += (fc ? "" : ",") + std::string("\"") + c.first + "\":" + std::to_string(c.second);
      fc = false;
    }
    js += "}";
  }
  return js + "}";
}

// DriverUtils.dumpGraph (DriverUtils.java:109-114): ">node_<key>\n" then the
// internal k-mer (or the key) and "\n", record by record.
static std::string dump_fasta(const uint8_t* recs, size_t n, int k) {
  std::string o;
  for (const GNode& x : parse_graph(recs, n, k)) {
    o += ">node_" + x.key.str() + "\n";
    o += (x.node.has_internal && x.node.internal.k > 0) ? x.node.internal.str() : x.key.str();
    o += "\n";
  }
  return o;
}

// FittingMixture.fittingMixture (FittingMixture.java:89-219), one data point
// at a time as the reference walks its expanded double[] (one entry per
// counted node, GraphStatistics.getCoverageStats, GraphStatistics.java:182-199).
// commons-math3 3.2 densities: ExponentialDistribution.density
// exp(-x/mean)/mean for x >= 0, else 0; NormalDistribution.density
// exp(-((x-mu)/sd)^2/2) / (sd sqrt(2 pi)).  out: cutoff, exponential mean,
// normal mean, normal sd, exponential prior, normal prior.
struct MixState {
  double exp_mean, normal_mean, normal_std;
};

static double ora_exp_pdf(double mean, double x) { return x < 0 ? 0.0 : std::exp(-x / mean) / mean; }
static double ora_normal_pdf(double mean, double sd, double x) {
  const double x1 = (x - mean) / sd;
  return std::exp(-0.5 * x1 * x1) / (sd * std::sqrt(2 * M_PI));
}

static void fitting_mixture(const double* data, size_t n, double max, int iterations, double* out) {
  MixState cur{5, 20, 5};  // :94-98
  double prob_exp = 0.5, prob_normal = 0.5;  // :102-103
  std::vector<double> m_exp(n), m_normal(n);
  double exp_sum = 0, normal_sum = 0;
  // :112-135 and :162-185: memberships with 1e-9 floors, normalised, then
  // the class priors from their sums.
  auto expectation = [&] {
    for (size_t i = 0; i < n; i++) {
      m_exp[i] = ora_exp_pdf(cur.exp_mean, data[i]) * prob_exp;
      if (m_exp[i] == 0) m_exp[i] = 0.000000001;
      m_normal[i] = ora_normal_pdf(cur.normal_mean, cur.normal_std, data[i]) * prob_normal;
      if (m_normal[i] == 0) m_normal[i] = 0.000000001;
      const double sum = m_exp[i] + m_normal[i];
      m_exp[i] = m_exp[i] / sum;
      m_normal[i] = m_normal[i] / sum;
    }
    exp_sum = 0;
    normal_sum = 0;
    for (size_t i = 0; i < n; i++) {
      exp_sum += m_exp[i];
      normal_sum += m_normal[i];
    }
    const double all = exp_sum + normal_sum;
    prob_exp = exp_sum / all;
    prob_normal = normal_sum / all;
  };
  expectation();
  for (int it = 0; it < iterations; it++) {  // :139-189
    cur.exp_mean = 0;
    cur.normal_mean = 0;
    cur.normal_std = 0;
    for (size_t i = 0; i < n; i++) {
      cur.exp_mean += m_exp[i] * data[i];
      cur.normal_mean += m_normal[i] * data[i];
    }
    cur.exp_mean = cur.exp_mean / exp_sum;
    if (cur.exp_mean <= 0) throw OraError(E_INVAL, "exponential mean must be positive");
    cur.normal_mean = cur.normal_mean / normal_sum;
    for (size_t i = 0; i < n; i++) cur.normal_std += m_normal[i] * std::pow(data[i] - cur.normal_mean, 2);
    cur.normal_std = std::sqrt(cur.normal_std / normal_sum);
    if (cur.normal_std == 0) cur.normal_std = 0.000000001f;
    expectation();
  }
  double cutoff = 0;  // :210-218
  for (double cov = 1; cov < max; cov++) {
    if (prob_exp * ora_exp_pdf(cur.exp_mean, cov) < prob_normal * ora_normal_pdf(cur.normal_mean, cur.normal_std, cov)) {
      cutoff = cov;
      break;
    }
  }
  out[0] = cutoff;
  out[1] = cur.exp_mean;
  out[2] = cur.normal_mean;
  out[3] = cur.normal_std;
  out[4] = prob_exp;
  out[5] = prob_normal;
}

}  // namespace ora

// ---------------------------------------------------------------------------
// C ABI for ctypes (oracle/oracle.py)
// ---------------------------------------------------------------------------

using namespace ora;

static thread_local std::string g_err;

struct ora_result {
  BuildResult r;
  std::vector<std::vector<uint8_t>> recs;
};

static int guard(const std::function<void()>& f) {
  try {
    f();
    return 0;
  } catch (const OraError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

static uint8_t* dup_bytes(const void* p, size_t n) {
  uint8_t* o = (uint8_t*)malloc(n ? n : 1);
  if (n) memcpy(o, p, n);
  return o;
}

extern "C" {

const char* ora_last_error() { return g_err.c_str(); }
// Test hook: 0 = JDK-7 HashSet (the model), 1/2 = planted wrong models.
void ora_set_hash_model(int m) { g_hash_model = m; }
void ora_free(void* p) { free(p); }

int ora_build(const char* text, size_t len, int k, int nparts, const int64_t* split_starts, int nsplits,
              int local_only, ora_result** out) {
  return guard([&] {
    auto* r = new ora_result;
    try {
      r->r = build((const uint8_t*)text, len, k, nparts, split_starts, nsplits, local_only != 0);
    } catch (...) {
      delete r;
      throw;
    }
    for (size_t p = 0; p < r->r.parts.size(); p++) r->recs.push_back(part_records(r->r, (int)p));
    *out = r;
  });
}

// build_threaded: P partitions, P producers over equal byte ranges of the
// input (line aligned), stages on `nthreads` threads.
int ora_build_threaded(const char* text, size_t len, int k, int nparts, int nthreads, ora_result** out) {
  return guard([&] {
    auto* r = new ora_result;
    try {
      r->r = build_threaded((const uint8_t*)text, len, k, nparts, nthreads);
    } catch (...) {
      delete r;
      throw;
    }
    for (size_t p = 0; p < r->r.parts.size(); p++) r->recs.push_back(part_records(r->r, (int)p));
    *out = r;
  });
}

int ora_result_nparts(const ora_result* r) { return (int)r->recs.size(); }
int64_t ora_result_nrecords(const ora_result* r, int part) { return (int64_t)r->r.parts[part].size(); }
size_t ora_result_part(const ora_result* r, int part, const uint8_t** data) {
  *data = r->recs[part].data();
  return r->recs[part].size();
}
void ora_result_free(ora_result* r) { delete r; }

int ora_seqfile(const uint8_t* recs, size_t n, const uint8_t* sync16, const char* keycls, const char* valcls,
                uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto o = seqfile_wrap(recs, n, sync16, keycls, valcls);
    *outlen = o.size();
    *out = dup_bytes(o.data(), o.size());
  });
}

int ora_crc_sidecar(const uint8_t* p, size_t n, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto o = crc_sidecar(p, n);
    *outlen = o.size();
    *out = dup_bytes(o.data(), o.size());
  });
}

int ora_dump_text(const uint8_t* recs, size_t n, char** out, size_t* outlen) {
  return guard([&] {
    auto s = dump_text(recs, n);
    *outlen = s.size();
    *out = (char*)dup_bytes(s.data(), s.size());
  });
}

int ora_fastq_to_readids(const char* m1, size_t n1, const char* m2, size_t n2, int paired, char** out,
                         size_t* outlen) {
  return guard([&] {
    auto s = fastq_to_readids((const uint8_t*)m1, n1, (const uint8_t*)m2, n2, paired != 0);
    *outlen = s.size();
    *out = (char*)dup_bytes(s.data(), s.size());
  });
}

// Primitive hooks for the tests (string-level in, string-level out).
int ora_pack(const char* s, int k, int reverse, uint8_t* out) {
  return guard([&] {
    size_t n = strlen(s);
    if (reverse) pack_reverse((const uint8_t*)s, n, 0, k, out);
    else pack_forward((const uint8_t*)s, n, 0, k, out);
  });
}
int ora_shift(uint8_t* b, int k, char c) {
  return guard([&] { shift_with_next(b, k, gene_code((uint8_t)c)); });
}
int ora_kmer_str(const uint8_t* b, int k, char* out) {
  return guard([&] {
    auto s = kmer_to_string(k, b, bytes_for_k(k));
    memcpy(out, s.c_str(), s.size() + 1);
  });
}
int32_t ora_vkmer_hash(const char* s) {
  auto v = VKmer::from_string((const uint8_t*)s, strlen(s));
  return v.hash_code();
}
// Union-update of two comma-separated k-mer lists (JDK-7 HashSet order);
// writes the comma-separated result.
int ora_union_update(const char* L, const char* M, char* out, size_t cap) {
  return guard([&] {
    auto split = [](const char* s) {
      std::vector<VKmer> v;
      std::string cur;
      for (const char* p = s;; p++) {
        if (*p == ',' || *p == 0) {
          if (!cur.empty()) v.push_back(VKmer::from_string((const uint8_t*)cur.data(), cur.size()));
          cur.clear();
          if (!*p) break;
        } else {
          cur.push_back(*p);
        }
      }
      return v;
    };
    auto r = union_update(split(L), split(M));
    std::string o;
    for (size_t i = 0; i < r.size(); i++) o += (i ? "," : "") + r[i].str();
    if (o.size() + 1 > cap) throw OraError(E_INVAL, "buffer too small");
    memcpy(out, o.c_str(), o.size() + 1);
  });
}
int64_t ora_make_uuid(int mate, int lib, int64_t readId, int offset) {
  int64_t v = 0;
  if (guard([&] { v = make_uuid(mate, lib, readId, offset); }) != 0) return INT64_MIN;
  return v;
}
// Summary of a merged record stream: records, total k-mers (unitig
// lengths), sum of coverage x length, and whether keys ascend in
// KmerPointable order.
int ora_unitig_summary(const uint8_t* recs, size_t n, int k, double* out4) {
  return guard([&] {
    size_t off = 0;
    double nrec = 0, nk = 0, cv = 0;
    std::vector<uint8_t> prev;
    bool asc = true;
    while (off < n) {
      if (off + 8 > n) throw OraError(E_INPUT, "truncated record stream");
      uint32_t recLen = get_be32(recs + off), keyLen = get_be32(recs + off + 4);
      if (off + 8 + recLen > n || keyLen > recLen) throw OraError(E_INPUT, "bad record");
      VKmer key;
      read_vkmer(recs + off + 8, keyLen, key);
      Node x;
      size_t used = x.read(recs + off + 8 + keyLen, recLen - keyLen);
      if (used != recLen - keyLen) throw OraError(E_INPUT, "trailing bytes in a Node");
      double m = x.has_internal ? (double)(x.internal.k - k + 1) : 1.0;
      nrec += 1;
      nk += m;
      cv += (double)x.cov * m;
      std::vector<uint8_t> rk(key.b.rbegin(), key.b.rend());
      if (!prev.empty() && !(prev < rk)) asc = false;
      prev = rk;
      off += 8 + recLen;
    }
    out4[0] = nrec; out4[1] = nk; out4[2] = cv; out4[3] = asc ? 1 : 0;
  });
}

int ora_clean_bad_coverage(const uint8_t* recs, size_t n, int k, float minc, float maxc, uint8_t** out,
                           size_t* outlen) {
  return guard([&] {
    auto r = clean_bad_coverage(recs, n, k, minc, maxc);
    *out = dup_bytes(r.data(), r.size());
    *outlen = r.size();
  });
}

int ora_clean_tips(const uint8_t* recs, size_t n, int k, int maxlen, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto r = clean_tips(recs, n, k, maxlen);
    *out = dup_bytes(r.data(), r.size());
    *outlen = r.size();
  });
}

int ora_clean_coverage_then_tips(const uint8_t* recs, size_t n, int k, float minc, float maxc, int maxlen,
                                 uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto r = clean_coverage_then_tips(recs, n, k, minc, maxc, maxlen);
    *out = dup_bytes(r.data(), r.size());
    *outlen = r.size();
  });
}

int ora_graph_stats(const uint8_t* recs, size_t n, int k, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto r = graph_stats(recs, n, k);
    *out = dup_bytes(r.data(), r.size());
    *outlen = r.size();
  });
}

int ora_dump_fasta(const uint8_t* recs, size_t n, int k, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto r = dump_fasta(recs, n, k);
    *out = dup_bytes(r.data(), r.size());
    *outlen = r.size();
  });
}

int ora_unitigs(const uint8_t* recs, size_t n, int k, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto o = unitigs(recs, n, k);
    *outlen = o.size();
    *out = dup_bytes(o.data(), o.size());
  });
}

int ora_unitigs_p4(const uint8_t* recs, size_t n, int k, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto o = unitigs_p4(recs, n, k);
    *outlen = o.size();
    *out = dup_bytes(o.data(), o.size());
  });
}

// Test hook for planted mistakes in the unitig restatements (0 = none).
void ora_set_unitig_plant(int m) { g_unitig_plant = m; }

// EDGETYPE.resolveEdgeThroughPath on type codes FF=0 FR=1 RF=2 RR=3;
// returns -1 (and sets the error) for an invalid path.
int ora_resolve_edge(int BtoA, int BtoC) {
  int r = -1;
  if (guard([&] {
        if (BtoA < 0 || BtoA > 3 || BtoC < 0 || BtoC > 3) throw OraError(E_INVAL, "bad edge type");
        r = resolve_edge_through_path(BtoA, BtoC);
      }) != 0)
    return -1;
  return r;
}

// VKmer.mergeWith{FF,FR,RF,RR}Kmer on letter strings.
int ora_vkmer_merge(int t, int initial, const char* self, const char* other, char* out, size_t cap) {
  return guard([&] {
    if (t < 0 || t > 3) throw OraError(E_INVAL, "bad edge type");
    auto s = vkmer_merge(t, initial, self, other);
    if (s.size() + 1 > cap) throw OraError(E_INVAL, "buffer too small");
    memcpy(out, s.c_str(), s.size() + 1);
  });
}

// Node.mergeEdges(t, other) on edge lists written "FF|FR|RF|RR", each list
// comma separated (an empty field = null list); writes self's lists back
// in the same form.
int ora_merge_edges(int t, const char* self, const char* other, char* out, size_t cap) {
  return guard([&] {
    if (t < 0 || t > 3) throw OraError(E_INVAL, "bad edge type");
    auto parse = [](const char* s) {
      Node nd;
      int field = 0;
      std::string cur;
      for (const char* p = s;; p++) {
        if (*p == ',' || *p == '|' || *p == 0) {
          if (!cur.empty()) {
            if (field > 3) throw OraError(E_INVAL, "more than four lists");
            nd.edges[field].push_back(VKmer::from_string((const uint8_t*)cur.data(), cur.size()));
            nd.edges_present[field] = true;
          }
          cur.clear();
          if (*p == '|') field++;
          if (!*p) break;
        } else {
          cur.push_back(*p);
        }
      }
      return nd;
    };
    Node a = parse(self);
    const Node b = parse(other);
    node_merge_edges(t, a, b);
    std::string o;
    for (int e = 0; e < 4; e++) {
      if (e) o += "|";
      for (size_t i = 0; i < a.edges[e].size(); i++) o += (i ? "," : "") + a.edges[e][i].str();
    }
    if (o.size() + 1 > cap) throw OraError(E_INVAL, "buffer too small");
    memcpy(out, o.c_str(), o.size() + 1);
  });
}

int ora_bubble_merge(const uint8_t* recs, size_t n, int k, float maxd, int maxlen, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto o = bubble_merge(recs, n, k, maxd, maxlen < 0 ? 5 * k : maxlen);
    *outlen = o.size();
    *out = dup_bytes(o.data(), o.size());
  });
}

int ora_clean_pipeline(const uint8_t* recs, size_t n, int k, float minc, float maxc, int tiplen, float maxd,
                       int maxlen, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto o = clean_pipeline(recs, n, k, minc, maxc, tiplen < 0 ? k : tiplen, maxd, maxlen < 0 ? 5 * k : maxlen);
    *outlen = o.size();
    *out = dup_bytes(o.data(), o.size());
  });
}

int ora_bubble_nodes(const uint8_t* recs, size_t n, int k, float maxd, int maxlen, uint8_t** out, size_t* outlen) {
  return guard([&] {
    auto o = bubble_nodes(recs, n, k, maxd, maxlen < 0 ? 5 * k : maxlen);
    *outlen = o.size();
    *out = dup_bytes(o.data(), o.size());
  });
}

int ora_fitting_mixture(const double* data, size_t n, double max, int iterations, double* out) {
  return guard([&] {
    // GenomixDriver.java:126-127
    if (max == 0 || n == 0) throw OraError(E_INVAL, "No information for coverage!");
    fitting_mixture(data, n, max, iterations, out);
  });
}

const char* ora_float_str(float f) {
  static thread_local std::string s;
  s = java_float_str(f);
  return s.c_str();
}

}  // extern "C"
