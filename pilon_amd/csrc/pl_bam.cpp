// pl_bam.cpp — BAM reader producing the engine's read batches (libpilonbam.so).
//
// Host C++ with libdeflate and std::thread.  The reading side of Pilon's BamFile
// (pilon/BamFile.scala:38-60,108-148) goes through htsjdk; what the pileup
// path takes from each SAMRecord is restated here from the SAMv1
// specification (BGZF section 4.1, BAM records section 4.2):
//
//   * a BGZF block is a gzip member whose extra field carries BC (BSIZE);
//     its CDATA is raw deflate, followed by CRC32 and ISIZE of the data;
//   * the BAM stream is "BAM\1", l_text, text, n_ref, {l_name, name, l_ref}
//     and then records of block_size + 32 fixed bytes + name, CIGAR, SEQ
//     (4-bit codes "=ACMGRSVTWYHKDBN", two per byte, first in the high
//     nibble), QUAL (0xFF when absent) and tags;
//   * a CIGAR of more than 65535 operations is stored as kSmN in the record
//     and in full in the CG:B,I tag.
//
// Each slice of the file (about 64 MiB of inflated BAM) goes through three
// steps, each in parallel: the blocks are inflated; the records are
// delimited (the block_size chain, walked per segment from a guessed start
// and checked against the previous segment's end) and given their output
// offsets (runs of one reference, their bases assigned in file order);
// and the records are decoded into the per-reference batches.  Non-ACGT
// reads get their nibbles copied afterwards, in file order.
#include "../../include/pilon_bam.h"

#include <libdeflate.h>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace {

thread_local std::string g_err;

pl_status fail(pl_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return s;
}

// BAM and BGZF integers are little-endian, as is the host
inline uint16_t rd16(const uint8_t* p) { uint16_t v; std::memcpy(&v, p, 2); return v; }
inline uint32_t rd32(const uint8_t* p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
inline int32_t rdi32(const uint8_t* p) { int32_t v; std::memcpy(&v, p, 4); return v; }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the end-of-file marker block of the SAMv1 specification (section 4.1.2)
const uint8_t kEofBlock[28] = {0x1f, 0x8b, 0x08, 0x04, 0x00, 0x00, 0x00, 0x00, 0x00, 0xff, 0x06, 0x00, 0x42, 0x43,
                               0x02, 0x00, 0x1b, 0x00, 0x03, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00};

constexpr size_t kSliceBytes = (size_t)64 << 20;   // inflated bytes per slice (PILON_BAM_SLICE_KIB overrides)

size_t slice_bytes() {
  const char* e = std::getenv("PILON_BAM_SLICE_KIB");
  const long v = e ? std::strtol(e, nullptr, 10) : 0;
  return v > 0 ? (size_t)v << 10 : kSliceBytes;
}

struct Block {
  size_t cdata, clen;      // raw deflate data within the compressed buffer
  uint32_t isize, crc;
  size_t uoff;             // destination within the slice
};

// 2-bit code of a BAM base nibble (A=1 C=2 G=4 T=8), -1 for any other code
const int8_t kNib2[16] = {-1, 0, 1, -1, 2, -1, -1, -1, 3, -1, -1, -1, -1, -1, -1, -1};

// the two bases of a SEQ byte (first base in the high nibble) as 4 bits,
// first base in the low 2; 0xFF when either is not A/C/G/T
struct PairTable {
  uint8_t v[256];
  PairTable() {
    for (int b = 0; b < 256; b++) {
      const int c0 = kNib2[b >> 4], c1 = kNib2[b & 15];
      v[b] = (c0 < 0 || c1 < 0) ? 0xFF : (uint8_t)(c0 | (c1 << 2));
    }
  }
};
const PairTable kPair;

// An array of a trivial type that only grows: its pages are anonymous
// memory, remapped (not copied) when it grows, and new elements are zero
// (fresh pages), first touched by the threads that write them.
template <class T>
struct Buf {
  T* p = nullptr;
  size_t n = 0, cap = 0, bytes = 0;
  Buf() = default;
  Buf(std::initializer_list<T> il) { for (const T& v : il) push_back(v); }
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  Buf(Buf&& o) noexcept : p(o.p), n(o.n), cap(o.cap), bytes(o.bytes) { o.p = nullptr; o.n = o.cap = o.bytes = 0; }
  Buf& operator=(Buf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p; n = o.n; cap = o.cap; bytes = o.bytes;
      o.p = nullptr; o.n = o.cap = o.bytes = 0;
    }
    return *this;
  }
  ~Buf() { release(); }
  void release() {
    if (p) munmap(p, bytes);
    p = nullptr;
    n = cap = bytes = 0;
  }
  void reserve(size_t c) {
    if (c <= cap) return;
    const size_t want = std::max(c, cap + cap / 2);
    const size_t nb = (want * sizeof(T) + 4095) / 4096 * 4096;
    void* q = p ? mremap(p, bytes, nb, MREMAP_MAYMOVE)
                : mmap(nullptr, nb, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
    if (q == MAP_FAILED) throw std::bad_alloc();
    p = (T*)q;
    bytes = nb;
    cap = nb / sizeof(T);
  }
  void resize(size_t m) { reserve(m); n = m; }
  void resize(size_t m, const T& v) {
    reserve(m);
    if (!(v == T())) for (size_t i = n; i < m; i++) p[i] = v;
    n = std::max(n, m);
  }
  void push_back(const T& v) { reserve(n + 1); p[n++] = v; }
  template <class It> void insert(T* at, It a, It b) {       // appends only
    (void)at;
    const size_t k = (size_t)(b - a);
    reserve(n + k);
    std::copy(a, b, p + n);
    n += k;
  }
  T* data() { return p; }
  const T* data() const { return p; }
  size_t size() const { return n; }
  bool empty() const { return n == 0; }
  T* begin() { return p; }
  T* end() { return p + n; }
  const T* begin() const { return p; }
  const T* end() const { return p + n; }
  T& back() { return p[n - 1]; }
  const T& back() const { return p[n - 1]; }
  T& operator[](size_t i) { return p[i]; }
  const T& operator[](size_t i) const { return p[i]; }
};

struct RefReads {
  Buf<int32_t> pos, tlen;
  Buf<uint16_t> flag;
  Buf<uint8_t> mapq;
  Buf<int64_t> cigar_off{0}, seq_off{0};
  Buf<uint32_t> cigar;
  Buf<uint8_t> bases2;
  Buf<int8_t> quals;
  std::vector<int64_t> exo_read, exo_off;
  std::vector<uint8_t> nib;
  int64_t n = 0;
  int32_t last_pos = INT32_MIN;
};

struct Rec {                // one placed record of the current slice
  size_t u;                 // record start (its block_size field) in the slice
  int32_t ref;
  int64_t i;                // index within the reference's batch
  int64_t coff, soff;       // CIGAR word and base offsets within the batch
  uint32_t ncig, lseq;
  size_t cig;               // CIGAR words in the slice (record body or CG tag)
};

template <class F>
void parallel_for(int nth, int64_t n, const F& f) {
  if (nth <= 1 || n < 2) { f(0, (int64_t)0, n); return; }
  const int t_n = (int)std::min<int64_t>(nth, n);
  std::vector<std::thread> th;
  th.reserve((size_t)t_n);
  for (int t = 0; t < t_n; t++) th.emplace_back([&f, t, t_n, n] { f(t, n * t / t_n, n * (t + 1) / t_n); });
  for (auto& x : th) x.join();
}

// The CG:B,I tag of a record's tag area [p, e): word count and first word
// offset, or false.  Tag value sizes follow SAMv1 section 4.2.4.
bool find_cg(const uint8_t* base, size_t p, size_t e, uint32_t* count, size_t* words) {
  while (p + 3 <= e) {
    const uint8_t t0 = base[p], t1 = base[p + 1], ty = base[p + 2];
    p += 3;
    size_t sz = 0;
    switch (ty) {
      case 'A': case 'c': case 'C': sz = 1; break;
      case 's': case 'S': sz = 2; break;
      case 'i': case 'I': case 'f': sz = 4; break;
      case 'Z': case 'H': {
        while (p < e && base[p]) p++;
        if (p >= e) return false;
        p++;
        continue;
      }
      case 'B': {
        if (p + 5 > e) return false;
        const uint8_t sub = base[p];
        const uint32_t cnt = rd32(base + p + 1);
        size_t es = 0;
        switch (sub) {
          case 'c': case 'C': es = 1; break;
          case 's': case 'S': es = 2; break;
          case 'i': case 'I': case 'f': es = 4; break;
          default: return false;
        }
        if (t0 == 'C' && t1 == 'G' && (sub == 'I' || sub == 'i')) {
          if (p + 5 + (size_t)cnt * 4 > e) return false;
          *count = cnt;
          *words = p + 5;
          return true;
        }
        sz = 5 + (size_t)cnt * es;
        break;
      }
      default: return false;
    }
    p += sz;
  }
  return false;
}

}  // namespace

// One reference of a BAI index (SAMv1 section 5.2): chunks [beg, end) of
// virtual file offsets (compressed block offset << 16 | offset in the
// inflated block) per bin, and the linear index of 16 kb windows.
struct BaiRef {
  std::vector<std::pair<uint32_t, std::vector<std::pair<uint64_t, uint64_t>>>> bins;
  std::vector<uint64_t> lin;
};

namespace {
// A file mapped read-only (no copy through a read buffer).
struct Mapped {
  const uint8_t* p = nullptr;
  size_t n = 0;
  Mapped() = default;
  Mapped(const Mapped&) = delete;
  Mapped& operator=(const Mapped&) = delete;
  ~Mapped() { if (p) munmap((void*)p, n); }
};
}  // namespace

struct pl_bam {
  std::vector<std::string> names;
  std::vector<int64_t> lens;
  std::vector<RefReads> refs;
  pl_bam_stats st;
  // opened through its index: records are read per query
  bool indexed = false;
  std::string path;
  int nth = 1;
  std::vector<BaiRef> bai;
  RefReads query;            // the last pl_bam_query's reads
  // opened as a stream (pl_bam_stream_open): the mapped file, the next
  // block, and the inflated bytes of the record the last slice cut
  bool streamed = false;
  Mapped map;
  size_t next_block = 0, last_at = SIZE_MAX, last_end = 0;
  Buf<uint8_t> carry;
  int64_t stream_bytes = 0;  // inflated bytes of every block (the ISIZE fields)
};

namespace {

// Header at the start of the inflated stream: bytes consumed, 0 if the
// stream does not hold all of it yet, -1 if it is not a BAM header.
int64_t parse_header(const uint8_t* u, size_t n, pl_bam* B) {
  if (n < 12) return 0;
  if (std::memcmp(u, "BAM\1", 4) != 0) return -1;
  const int32_t l_text = rdi32(u + 4);
  if (l_text < 0) return -1;
  size_t p = 8 + (size_t)l_text;
  if (p + 4 > n) return 0;
  const int32_t n_ref = rdi32(u + p);
  if (n_ref < 0) return -1;
  p += 4;
  std::vector<std::string> names;
  std::vector<int64_t> lens;
  for (int32_t r = 0; r < n_ref; r++) {
    if (p + 4 > n) return 0;
    const int32_t l_name = rdi32(u + p);
    if (l_name < 1) return -1;
    if (p + 4 + (size_t)l_name + 4 > n) return 0;
    names.emplace_back((const char*)u + p + 4, strnlen((const char*)u + p + 4, (size_t)l_name));
    lens.push_back(rdi32(u + p + 4 + l_name));
    p += 8 + (size_t)l_name;
  }
  B->names = std::move(names);
  B->lens = std::move(lens);
  B->refs.resize((size_t)n_ref);
  return (int64_t)p;
}

// Inflate blocks of the compressed bytes cb into out (block b at out + b.uoff).
pl_status inflate_blocks(const uint8_t* cb, std::vector<Block>& blocks, uint8_t* out, int nth) {
  std::atomic<int64_t> bad{-1};
  parallel_for(nth, (int64_t)blocks.size(), [&](int, int64_t lo, int64_t hi) {
    libdeflate_decompressor* dz = libdeflate_alloc_decompressor();
    if (!dz) { bad = lo; return; }
    for (int64_t k = lo; k < hi; k++) {
      const Block& b = blocks[(size_t)k];
      uint8_t none = 0;
      uint8_t* dst = b.isize ? out + b.uoff : &none;   // an empty block (the EOF marker) still inflates
      size_t got = 0;
      const libdeflate_result rc = libdeflate_deflate_decompress(dz, cb + b.cdata, b.clen, dst,
                                                                 b.isize ? b.isize : 1, &got);
      if (rc != LIBDEFLATE_SUCCESS || got != b.isize || libdeflate_crc32(0, dst, b.isize) != b.crc) {
        int64_t e = -1;
        bad.compare_exchange_strong(e, k);
      }
    }
    libdeflate_free_decompressor(dz);
  });
  if (bad >= 0) return fail(PL_E_INVAL, "corrupt BGZF block (inflate, ISIZE or CRC32 mismatch)");
  return PL_OK;
}

// A record start that fits the SAMv1 record layout (block_size, refIDs,
// a NUL-terminated name, the variable fields within block_size): the test
// a speculative record search uses.
bool plausible_record(const uint8_t* u, size_t q, size_t n, int32_t nref) {
  if (q + 36 > n) return false;
  const uint8_t* r = u + q;
  const uint32_t bs = rd32(r);
  if (bs < 32 || bs > (1u << 28)) return false;
  const int32_t ref = rdi32(r + 4), pos = rdi32(r + 8), next_ref = rdi32(r + 24);
  if (ref < -1 || ref >= nref || next_ref < -1 || next_ref >= nref || pos < -1) return false;
  const uint32_t l_name = r[12], ncig = rd16(r + 16), lseq = rd32(r + 20);
  if (l_name < 1 || lseq > bs) return false;
  if (36 + (size_t)l_name + 4 * (size_t)ncig + ((size_t)lseq + 1) / 2 + lseq > 4 + (size_t)bs) return false;
  return q + 36 + l_name > n || r[36 + l_name - 1] == 0;
}

// The block_size chain from record start p: the starts below `end` go to
// offs; *exit is the first start at or past `end`, or where the data stops
// holding a whole record; *bad is the start of a record whose block_size is
// below the fixed 32 bytes (the chain stops there).
void walk_chain(const uint8_t* u, size_t p, size_t end, size_t n, std::vector<size_t>& offs, size_t* exit,
                size_t* bad) {
  *bad = SIZE_MAX;
  while (p < end && p + 4 <= n) {
    const uint32_t bs = rd32(u + p);
    if (bs < 32) { *bad = p; break; }
    if (p + 4 + (size_t)bs > n) break;
    offs.push_back(p);
    p += 4 + (size_t)bs;
  }
  *exit = p;
}

// Record starts of ub[u0, n), as the sequential block_size chain gives
// them, found in parallel: the range is cut into one segment per thread;
// each segment after the first guesses its first record start (the first
// offset from which three records in a row fit the layout) and walks the
// chain from there; a pass over the segments in order then keeps a walk
// only where its guess is the previous segment's exit, and walks again
// from that exit where it is not.  The result is the sequential chain
// whatever the guesses.  *used = the start of the first incomplete record.
pl_status delimit_records(pl_bam* B, const uint8_t* u, size_t u0, size_t n, int nth, std::vector<size_t>& offs,
                          size_t* used) {
  const int32_t nref = (int32_t)B->refs.size();
  const int T = (n - u0 > ((size_t)4 << 20)) ? std::max(1, nth) : 1;
  struct Seg { size_t lo, hi, start = SIZE_MAX, exit = 0, bad = SIZE_MAX; std::vector<size_t> offs; };
  std::vector<Seg> seg((size_t)T);
  for (int t = 0; t < T; t++) {
    seg[(size_t)t].lo = u0 + (n - u0) * (size_t)t / (size_t)T;
    seg[(size_t)t].hi = t + 1 == T ? n : u0 + (n - u0) * (size_t)(t + 1) / (size_t)T;
  }
  parallel_for(T, T, [&](int, int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; t++) {
      Seg& S = seg[(size_t)t];
      if (t == 0) {
        S.start = u0;
      } else {
        for (size_t q = S.lo; q < S.hi && q < S.lo + ((size_t)1 << 20); q++) {
          bool ok = true;
          size_t r = q;
          for (int k = 0; k < 3 && ok && r + 36 <= n; k++) {
            ok = plausible_record(u, r, n, nref);
            r += 4 + (size_t)rd32(u + r);
          }
          if (ok) { S.start = q; break; }
        }
      }
      if (S.start != SIZE_MAX) {
        S.offs.reserve((S.hi - S.lo) / 200 + 16);
        walk_chain(u, S.start, S.hi, n, S.offs, &S.exit, &S.bad);
      }
    }
  });
  size_t expected = u0;
  for (int t = 0; t < T; t++) {
    Seg& S = seg[(size_t)t];
    if (S.start != expected) {                 // a wrong guess: the chain from the true start
      S.offs.clear();
      walk_chain(u, expected, t + 1 == T ? n : std::max(S.hi, expected), n, S.offs, &S.exit, &S.bad);
    }
    if (S.bad != SIZE_MAX)
      return fail(PL_E_INVAL, "BAM record %lld: block_size %u below the fixed 32 bytes",
                  (long long)(B->st.records + (int64_t)offs.size() + (int64_t)S.offs.size()), rd32(u + S.bad));
    offs.insert(offs.end(), S.offs.begin(), S.offs.end());
    expected = S.exit;
  }
  *used = expected;
  return PL_OK;
}

// Delimit the records of ub[u0, n) and decode every placed one into its
// reference's batch; returns the offset of the first incomplete record.
pl_status decode_records(pl_bam* B, const uint8_t* u, size_t u0, size_t n, int nth, size_t* used) {
  const int64_t nref = (int64_t)B->refs.size();
  std::vector<size_t> offs;
  if (pl_status s = delimit_records(B, u, u0, n, nth, offs, used)) return s;
  // per-record facts, in parallel: lengths checked, the CG tag resolved
  const int64_t N = (int64_t)offs.size();
  std::vector<Rec> all((size_t)N);
  std::atomic<int64_t> bad{INT64_MAX};
  parallel_for(nth, N, [&](int, int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; k++) {
      const size_t q = offs[(size_t)k];
      const uint8_t* r = u + q;
      const uint32_t bs = rd32(r);
      const uint32_t l_name = r[12];
      uint32_t ncig = rd16(r + 16);
      const uint32_t lseq = rd32(r + 20);
      const size_t body = 36 + (size_t)l_name + 4 * (size_t)ncig;
      if (lseq > bs || body + ((size_t)lseq + 1) / 2 + lseq > 4 + (size_t)bs) {
        int64_t e = bad.load();
        while (k < e && !bad.compare_exchange_weak(e, k)) {}
        continue;
      }
      Rec& rc = all[(size_t)k];
      rc.u = q;
      rc.ref = rdi32(r + 4);
      rc.i = rdi32(r + 8);            // pos (0-based) until the offsets are assigned
      size_t cig = q + 36 + l_name;
      // kSmN placeholder: the real CIGAR is in the CG:B,I tag
      if (ncig == 2 && rd32(u + cig) == ((lseq << 4) | 4u) && (rd32(u + cig + 4) & 15u) == 3u) {
        uint32_t cnt = 0;
        size_t words = 0;
        if (find_cg(u, q + body + ((size_t)lseq + 1) / 2 + lseq, q + 4 + bs, &cnt, &words)) { ncig = cnt; cig = words; }
      }
      rc.ncig = ncig;
      rc.lseq = lseq;
      rc.cig = cig;
    }
  });
  if (bad.load() != INT64_MAX)
    return fail(PL_E_INVAL, "BAM record %lld: name, CIGAR, SEQ and QUAL exceed block_size",
                (long long)(B->st.records + bad.load()));
  // output offsets per reference, in file order: runs of one reference per
  // chunk (in parallel), their bases assigned in file order (serially, one
  // step per run), then every record's offsets (in parallel)
  struct Run {
    int32_t ref;
    int64_t k0, k1;                    // records [k0, k1)
    int64_t nrec = 0, ncig = 0, nseq = 0;
    int32_t first_pos, last_pos;
    int64_t unsorted = -1;             // first record placed before its predecessor
    int64_t i0 = 0, c0 = 0, s0 = 0;
  };
  const int TC = (int)std::max<int64_t>(1, std::min<int64_t>(nth, N / 4096 + 1));
  std::vector<std::vector<Run>> runs((size_t)TC);
  std::vector<int64_t> badref((size_t)TC, -1);
  parallel_for(TC, TC, [&](int, int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; t++) {
      const int64_t a = N * t / TC, b = N * (t + 1) / TC;
      std::vector<Run>& V = runs[(size_t)t];
      for (int64_t k = a; k < b; k++) {
        const Rec& rc = all[(size_t)k];
        if (rc.ref >= nref || rc.ref < -1) { badref[(size_t)t] = k; break; }
        if (rc.ref < 0) continue;
        const int32_t pos1 = (int32_t)rc.i + 1;
        if (V.empty() || V.back().ref != rc.ref || V.back().k1 != k) {
          Run r;
          r.ref = rc.ref; r.k0 = k; r.k1 = k; r.first_pos = pos1; r.last_pos = pos1;
          V.push_back(r);
        }
        Run& r = V.back();
        if (pos1 < r.last_pos && r.unsorted < 0) r.unsorted = k;
        r.last_pos = std::max(r.last_pos, pos1);
        r.k1 = k + 1;
        r.nrec++;
        r.ncig += rc.ncig;
        r.nseq += rc.lseq;
      }
    }
  });
  std::vector<int64_t> add_n((size_t)nref, 0), add_c((size_t)nref, 0), add_s((size_t)nref, 0);
  std::vector<int32_t> touched;
  int64_t placed = 0;
  for (int t = 0; t < TC; t++) {
    for (Run& r : runs[(size_t)t]) {
      RefReads& R = B->refs[(size_t)r.ref];
      if (r.first_pos < R.last_pos || r.unsorted >= 0) {
        // the first record in file order placed before its predecessor
        int32_t prev = R.last_pos;
        for (int64_t k = r.k0; k < r.k1; k++) {
          const Rec& rc = all[(size_t)k];
          if (rc.ref != r.ref) continue;
          const int32_t pos1 = (int32_t)rc.i + 1;
          if (pos1 < prev)
            return fail(PL_E_INVAL, "BAM is not coordinate-sorted: %s pos %d after %d", B->names[(size_t)r.ref].c_str(),
                        pos1, prev);
          prev = pos1;
        }
      }
      R.last_pos = r.last_pos;
      if (add_n[(size_t)r.ref] == 0) touched.push_back(r.ref);
      r.i0 = R.n + add_n[(size_t)r.ref];
      r.c0 = R.cigar_off.back() + add_c[(size_t)r.ref];
      r.s0 = R.seq_off.back() + add_s[(size_t)r.ref];
      add_n[(size_t)r.ref] += r.nrec;
      add_c[(size_t)r.ref] += r.ncig;
      add_s[(size_t)r.ref] += r.nseq;
      placed += r.nrec;
    }
    if (badref[(size_t)t] >= 0) {
      const int64_t k = badref[(size_t)t];
      return fail(PL_E_INVAL, "BAM record %lld: refID %d outside the dictionary", (long long)(B->st.records + k),
                  all[(size_t)k].ref);
    }
  }
  parallel_for(TC, TC, [&](int, int64_t lo, int64_t hi) {
    for (int64_t t = lo; t < hi; t++)
      for (const Run& r : runs[(size_t)t]) {
        int64_t i = r.i0, c = r.c0, so = r.s0;
        for (int64_t k = r.k0; k < r.k1; k++) {
          Rec& rc = all[(size_t)k];
          if (rc.ref != r.ref) continue;
          rc.i = i++;
          rc.coff = c;
          rc.soff = so;
          c += rc.ncig;
          so += rc.lseq;
        }
      }
  });
  B->st.records += N;
  // grow the batches of the references this slice touches (new bytes are
  // zero: bytes of bases2 that two reads share are OR-ed by both)
  for (int32_t ref : touched) {
    RefReads& R = B->refs[(size_t)ref];
    const size_t nn = (size_t)(R.n + add_n[(size_t)ref]);
    const int64_t cn = R.cigar_off.back() + add_c[(size_t)ref];
    const int64_t sn = R.seq_off.back() + add_s[(size_t)ref];
    R.pos.resize(nn); R.tlen.resize(nn); R.flag.resize(nn); R.mapq.resize(nn);
    R.cigar_off.resize(nn + 1); R.seq_off.resize(nn + 1);
    R.cigar.resize((size_t)cn);
    R.quals.resize((size_t)sn);
    R.bases2.resize((size_t)(sn + 3) / 4);
  }
  // decode in parallel; bytes of bases2 shared by two reads are OR-ed atomically
  std::vector<uint8_t> exotic((size_t)N, 0);
  parallel_for(nth, N, [&](int, int64_t lo, int64_t hi) {
    for (int64_t k = lo; k < hi; k++) {
      const Rec& rc = all[(size_t)k];
      if (rc.ref < 0) continue;
      RefReads& R = B->refs[(size_t)rc.ref];
      const uint8_t* r = u + rc.u;
      const uint32_t l_name = r[12];
      const uint16_t flag = rd16(r + 18);
      const int32_t next_ref = rdi32(r + 24);
      const size_t i = (size_t)rc.i;
      R.pos[i] = rdi32(r + 8) + 1;
      R.tlen[i] = rdi32(r + 32);
      R.mapq[i] = r[13];
      R.cigar_off[i + 1] = rc.coff + rc.ncig;
      R.seq_off[i + 1] = rc.soff + rc.lseq;
      std::memcpy(R.cigar.data() + rc.coff, u + rc.cig, 4 * (size_t)rc.ncig);
      const uint8_t* seq = r + 36 + l_name + 4 * (size_t)rd16(r + 16);
      const uint8_t* qual = seq + ((size_t)rc.lseq + 1) / 2;
      const bool noqual = rc.lseq == 0 || qual[0] == 0xFF;
      std::memcpy(R.quals.data() + rc.soff, qual, rc.lseq);
      // two bases per SEQ byte through kPair: any code but A/C/G/T makes the
      // read exotic (the padding nibble of an odd length is not a base)
      const uint32_t npair = rc.lseq >> 1;
      uint8_t orp = 0;
      for (uint32_t j = 0; j < npair; j++) orp |= kPair.v[seq[j]] == 0xFF ? 1 : 0;
      bool exo = orp != 0 || ((rc.lseq & 1) && kNib2[seq[npair] >> 4] < 0);
      if (!exo && rc.lseq) {
        // 2-bit codes, four to a byte, emitted a byte at a time; the first
        // and last byte may hold bases of the neighbouring reads
        const int64_t s0 = rc.soff, s1 = rc.soff + rc.lseq;
        uint8_t* out = R.bases2.data();
        int64_t byte = s0 >> 2;
        uint32_t acc = 0;
        int bits = 2 * (int)(s0 & 3);
        auto emit = [&](uint8_t v) {
          const bool whole = (byte << 2) >= s0 && (byte << 2) + 4 <= s1;
          if (whole) out[byte] = v;
          else __atomic_fetch_or(&out[byte], v, __ATOMIC_RELAXED);
          byte++;
        };
        for (uint32_t j = 0; j < npair; j++) {
          acc |= (uint32_t)kPair.v[seq[j]] << bits;
          bits += 4;
          if (bits >= 8) { emit((uint8_t)acc); acc >>= 8; bits -= 8; }
        }
        if (rc.lseq & 1) { acc |= (uint32_t)kNib2[seq[npair] >> 4] << bits; bits += 2; }
        while (bits > 0) { emit((uint8_t)acc); acc >>= 8; bits -= 8; }
      }
      exotic[(size_t)k] = exo ? 1 : 0;
      R.flag[i] = (uint16_t)((flag & 0x0FFFu) | (noqual ? PL_RF_NO_QUAL : 0u) |
                             (next_ref == rc.ref ? PL_RF_MATE_SAME_REF : 0u) | (exo ? PL_RF_EXOTIC : 0u));
    }
  });
  // nibbles of the non-ACGT reads, in file order, each from an even nibble
  for (size_t k = 0; k < (size_t)N; k++) {
    if (!exotic[k]) continue;
    const Rec& rc = all[k];
    RefReads& R = B->refs[(size_t)rc.ref];
    const uint8_t* r = u + rc.u;
    const uint8_t* seq = r + 36 + r[12] + 4 * (size_t)rd16(r + 16);
    R.exo_read.push_back(rc.i);
    R.exo_off.push_back((int64_t)R.nib.size() * 2);
    R.nib.insert(R.nib.end(), seq, seq + ((size_t)rc.lseq + 1) / 2);
    if (rc.lseq & 1) R.nib.back() &= 0xF0;
    B->st.exotic++;
  }
  for (int32_t ref : touched) B->refs[(size_t)ref].n += add_n[(size_t)ref];
  B->st.placed += placed;
  return PL_OK;
}

pl_status map_file(const char* path, Mapped& m) {
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return fail(PL_E_INVAL, "cannot open %s", path);
  struct stat sb;
  std::memset(&sb, 0, sizeof sb);
  if (fstat(fd, &sb) == 0 && sb.st_size > 0) {
    void* q = mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE | MAP_POPULATE, fd, 0);
    if (q != MAP_FAILED) {
      m.p = (const uint8_t*)q;
      m.n = (size_t)sb.st_size;
      (void)madvise(q, m.n, MADV_SEQUENTIAL);
    }
  }
  ::close(fd);
  if (sb.st_size > 0 && !m.p) return fail(PL_E_INVAL, "cannot map %s", path);
  return PL_OK;
}

// The whole BGZF blocks of the mapped file from byte q whose inflated bytes
// fit `room` (at least one block when any is there); *end = the byte after
// the last one taken.
pl_status next_blocks(const Mapped& m, size_t q, size_t room, std::vector<Block>& blocks, size_t* end, size_t* last_at) {
  const uint8_t* cb = m.p;
  size_t utot = 0;
  blocks.clear();
  while (q + 18 <= m.n) {
    const uint8_t* h = cb + q;
    if (h[0] != 0x1f || h[1] != 0x8b || h[2] != 8 || !(h[3] & 4))
      return fail(PL_E_INVAL, "not a BGZF block at byte %lld", (long long)q);
    const size_t xlen = rd16(h + 10);
    if (q + 12 + xlen > m.n) break;
    size_t bsize = 0;
    for (size_t x = 12; x + 4 <= 12 + xlen;) {
      const size_t slen = rd16(h + x + 2);
      if (h[x] == 66 && h[x + 1] == 67 && slen == 2) bsize = (size_t)rd16(h + x + 4) + 1;
      x += 4 + slen;
    }
    if (!bsize || bsize < 20 + xlen) return fail(PL_E_INVAL, "BGZF block without a valid BC (BSIZE) subfield");
    if (q + bsize > m.n) break;
    Block bk;
    bk.cdata = q + 12 + xlen;
    bk.clen = bsize - xlen - 20;
    bk.crc = rd32(h + bsize - 8);
    bk.isize = rd32(h + bsize - 4);
    if (bk.isize > 65536) return fail(PL_E_INVAL, "BGZF block ISIZE %u above 64 KiB", bk.isize);
    if (!blocks.empty() && utot + bk.isize > room) break;
    bk.uoff = utot;
    utot += bk.isize;
    blocks.push_back(bk);
    *last_at = q;
    q += bsize;
  }
  if (blocks.empty() && q < m.n) return fail(PL_E_INVAL, "truncated BGZF block at the end of the file");
  *end = q;
  return PL_OK;
}

pl_status open_impl(const char* path, int nth, pl_bam* B) {
  Mapped m;
  if (pl_status s = map_file(path, m)) return s;
  const uint8_t* cb = m.p;       // the compressed file
  B->st.compressed_bytes = (int64_t)m.n;
  Buf<uint8_t> ub;               // inflated bytes not yet decoded
  size_t carry = 0, p = 0, last_at = SIZE_MAX, last_end = 0;
  const size_t slice = slice_bytes();
  bool header = false;
  while (p < m.n) {
    // whole blocks from p, up to one slice of inflated bytes
    std::vector<Block> blocks;
    size_t q = p, utot = 0;
    while (q + 18 <= m.n && utot < slice) {
      const uint8_t* h = cb + q;
      if (h[0] != 0x1f || h[1] != 0x8b || h[2] != 8 || !(h[3] & 4))
        return fail(PL_E_INVAL, "not a BGZF block at byte %lld", (long long)q);
      const size_t xlen = rd16(h + 10);
      if (q + 12 + xlen > m.n) break;
      size_t bsize = 0;
      for (size_t x = 12; x + 4 <= 12 + xlen;) {
        const size_t slen = rd16(h + x + 2);
        if (h[x] == 66 && h[x + 1] == 67 && slen == 2) bsize = (size_t)rd16(h + x + 4) + 1;
        x += 4 + slen;
      }
      if (!bsize || bsize < 20 + xlen) return fail(PL_E_INVAL, "BGZF block without a valid BC (BSIZE) subfield");
      if (q + bsize > m.n) break;
      Block bk;
      bk.cdata = q + 12 + xlen;
      bk.clen = bsize - xlen - 20;
      bk.crc = rd32(h + bsize - 8);
      bk.isize = rd32(h + bsize - 4);
      bk.uoff = utot;
      if (bk.isize > 65536) return fail(PL_E_INVAL, "BGZF block ISIZE %u above 64 KiB", bk.isize);
      utot += bk.isize;
      blocks.push_back(bk);
      last_at = q;
      last_end = q + bsize;
      q += bsize;
    }
    if (blocks.empty()) return fail(PL_E_INVAL, "truncated BGZF block at the end of the file");
    B->st.blocks += (int64_t)blocks.size();
    ub.resize(carry + utot);
    double t0 = now_s();
    pl_status s = inflate_blocks(cb, blocks, ub.data() + carry, nth);
    B->st.inflate_s += now_s() - t0;
    if (s) return s;
    B->st.bam_bytes += (int64_t)utot;
    p = q;
    size_t un = carry + utot, u0 = 0;
    if (!header) {
      const int64_t h = parse_header(ub.data(), un, B);
      if (h < 0) return fail(PL_E_INVAL, "not a BAM file (bad magic or header)");
      if (h == 0) { carry = un; continue; }
      header = true;
      u0 = (size_t)h;
    }
    t0 = now_s();
    size_t used = u0;
    s = decode_records(B, ub.data(), u0, un, nth, &used);
    B->st.decode_s += now_s() - t0;
    if (s) return s;
    carry = un - used;
    if (carry) std::memmove(ub.data(), ub.data() + used, carry);
  }
  if (!header) return fail(PL_E_INVAL, "not a BAM file (no complete header)");
  if (carry) return fail(PL_E_INVAL, "truncated BAM record at the end of the file");
  B->st.eof_marker = last_at != SIZE_MAX && last_end - last_at == sizeof kEofBlock &&
                     std::memcmp(cb + last_at, kEofBlock, sizeof kEofBlock) == 0;
  return PL_OK;
}

// pl_bam_stream_next's work: the carried bytes, then whole blocks that fit
// the caller's buffer, inflated there; the records delimited; the cut
// record carried to the next call.
pl_status stream_next_impl(pl_bam* B, uint8_t* dst, int64_t cap, int64_t* len, int64_t* rec_off, int64_t rec_cap,
                           int64_t* n_rec) {
  *len = 0;
  *n_rec = 0;
  const size_t cc = B->carry.size();
  if ((size_t)cap < cc + 65536) return fail(PL_E_INVAL, "slice buffer of %lld bytes below one record and one block", (long long)cap);
  if (cc) std::memcpy(dst, B->carry.data(), cc);
  std::vector<Block> blocks;
  size_t end = B->next_block;
  if (pl_status s = next_blocks(B->map, B->next_block, (size_t)cap - cc, blocks, &end, &B->last_at)) return s;
  size_t utot = 0;
  for (const Block& b : blocks) utot += b.isize;
  if (blocks.empty()) {
    if (cc) return fail(PL_E_INVAL, "truncated BAM record at the end of the file");
    B->st.eof_marker = B->last_at != SIZE_MAX && B->last_end - B->last_at == sizeof kEofBlock &&
                       std::memcmp(B->map.p + B->last_at, kEofBlock, sizeof kEofBlock) == 0;
    return PL_OK;
  }
  B->last_end = end;
  B->st.blocks += (int64_t)blocks.size();
  double t0 = now_s();
  if (pl_status s = inflate_blocks(B->map.p, blocks, dst + cc, B->nth)) return s;
  B->st.inflate_s += now_s() - t0;
  B->st.bam_bytes += (int64_t)utot;
  B->next_block = end;
  const size_t un = cc + utot;
  t0 = now_s();
  std::vector<size_t> offs;
  size_t used = 0;
  if (pl_status s = delimit_records(B, dst, 0, un, B->nth, offs, &used)) return s;
  B->st.decode_s += now_s() - t0;
  if ((int64_t)offs.size() > rec_cap)
    return fail(PL_E_INVAL, "%zu records in a slice above the %lld offsets given", offs.size(), (long long)rec_cap);
  for (size_t k = 0; k < offs.size(); k++) rec_off[k] = (int64_t)offs[k];
  B->st.records += (int64_t)offs.size();
  B->carry.resize(0);
  B->carry.resize(un - used);
  if (un > used) std::memcpy(B->carry.data(), dst + used, un - used);
  *len = (int64_t)used;
  *n_rec = (int64_t)offs.size();
  return PL_OK;
}

// The bins that can hold records overlapping [beg, end), 0-based (SAMv1
// section 5.3, reg2bins).
std::vector<uint32_t> reg2bins(int64_t beg, int64_t end) {
  std::vector<uint32_t> v{0};
  if (end <= beg) return v;
  end -= 1;
  for (int64_t k = 1 + (beg >> 26); k <= 1 + (end >> 26); k++) v.push_back((uint32_t)k);
  for (int64_t k = 9 + (beg >> 23); k <= 9 + (end >> 23); k++) v.push_back((uint32_t)k);
  for (int64_t k = 73 + (beg >> 20); k <= 73 + (end >> 20); k++) v.push_back((uint32_t)k);
  for (int64_t k = 585 + (beg >> 17); k <= 585 + (end >> 17); k++) v.push_back((uint32_t)k);
  for (int64_t k = 4681 + (beg >> 14); k <= 4681 + (end >> 14); k++) v.push_back((uint32_t)k);
  return v;
}

constexpr uint32_t kBaiMetaBin = 37450;   // pseudo-bin of per-reference counts, not a bin of records

// The BAI next to a BAM: path + ".bai", or path with ".bam" replaced by ".bai"
// (htsjdk's two names), or the caller's.
pl_status read_bai(const std::string& path, size_t n_ref, std::vector<BaiRef>& out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return fail(PL_E_INVAL, "cannot open index %s", path.c_str());
  std::unique_ptr<FILE, int (*)(FILE*)> fh(f, std::fclose);
  std::vector<uint8_t> d;
  uint8_t buf[1 << 16];
  for (size_t k; (k = std::fread(buf, 1, sizeof buf, f)) > 0;) d.insert(d.end(), buf, buf + k);
  size_t p = 0;
  auto need = [&](size_t n) { return p + n <= d.size(); };
  if (!need(8) || std::memcmp(d.data(), "BAI\1", 4) != 0) return fail(PL_E_INVAL, "%s is not a BAI index", path.c_str());
  const int32_t nr = rdi32(d.data() + 4);
  p = 8;
  if (nr < 0 || (size_t)nr != n_ref)
    return fail(PL_E_INVAL, "index %s has %d references, the BAM %zu", path.c_str(), nr, n_ref);
  out.assign((size_t)nr, BaiRef());
  for (int32_t r = 0; r < nr; r++) {
    if (!need(4)) return fail(PL_E_INVAL, "truncated index %s", path.c_str());
    const int32_t nbin = rdi32(d.data() + p);
    p += 4;
    if (nbin < 0) return fail(PL_E_INVAL, "corrupt index %s", path.c_str());
    for (int32_t b = 0; b < nbin; b++) {
      if (!need(8)) return fail(PL_E_INVAL, "truncated index %s", path.c_str());
      const uint32_t bin = rd32(d.data() + p);
      const int32_t nch = rdi32(d.data() + p + 4);
      p += 8;
      if (nch < 0 || !need(16 * (size_t)nch)) return fail(PL_E_INVAL, "truncated index %s", path.c_str());
      std::vector<std::pair<uint64_t, uint64_t>> ch((size_t)nch);
      for (int32_t c = 0; c < nch; c++) {
        uint64_t a, e;
        std::memcpy(&a, d.data() + p, 8);
        std::memcpy(&e, d.data() + p + 8, 8);
        ch[(size_t)c] = {a, e};
        p += 16;
      }
      if (bin != kBaiMetaBin) out[(size_t)r].bins.emplace_back(bin, std::move(ch));
    }
    if (!need(4)) return fail(PL_E_INVAL, "truncated index %s", path.c_str());
    const int32_t ni = rdi32(d.data() + p);
    p += 4;
    if (ni < 0 || !need(8 * (size_t)ni)) return fail(PL_E_INVAL, "truncated index %s", path.c_str());
    out[(size_t)r].lin.resize((size_t)ni);
    if (ni) std::memcpy(out[(size_t)r].lin.data(), d.data() + p, 8 * (size_t)ni);
    p += 8 * (size_t)ni;
    std::sort(out[(size_t)r].bins.begin(), out[(size_t)r].bins.end(),
              [](const std::pair<uint32_t, std::vector<std::pair<uint64_t, uint64_t>>>& a,
                 const std::pair<uint32_t, std::vector<std::pair<uint64_t, uint64_t>>>& b) { return a.first < b.first; });
  }
  return PL_OK;
}

// Whole BGZF blocks of the file from compressed offset `at`: reads until the
// block starting at `last` (inclusive) is in, or the file ends.  Returns the
// compressed bytes and the blocks within them.
pl_status read_blocks(FILE* f, uint64_t at, uint64_t last, std::vector<uint8_t>& cb, std::vector<Block>& blocks,
                      std::vector<uint64_t>& coff) {
  cb.clear();
  blocks.clear();
  coff.clear();
  if (fseeko(f, (off_t)at, SEEK_SET) != 0) return fail(PL_E_INVAL, "cannot seek to %llu", (unsigned long long)at);
  size_t utot = 0;
  for (uint64_t c = at; c <= last;) {
    uint8_t h[18];
    if (std::fread(h, 1, 18, f) != 18) break;                     // end of file
    if (h[0] != 0x1f || h[1] != 0x8b || h[2] != 8 || !(h[3] & 4))
      return fail(PL_E_INVAL, "not a BGZF block at byte %llu", (unsigned long long)c);
    const size_t xlen = rd16(h + 10);
    std::vector<uint8_t> blk(h, h + 18);
    blk.resize(12 + xlen);
    if (xlen > 6 && std::fread(blk.data() + 18, 1, xlen - 6, f) != xlen - 6)
      return fail(PL_E_INVAL, "truncated BGZF block at byte %llu", (unsigned long long)c);
    size_t bsize = 0;
    for (size_t x = 12; x + 4 <= 12 + xlen;) {
      const size_t slen = rd16(blk.data() + x + 2);
      if (blk[x] == 66 && blk[x + 1] == 67 && slen == 2) bsize = (size_t)rd16(blk.data() + x + 4) + 1;
      x += 4 + slen;
    }
    if (!bsize || bsize < 20 + xlen) return fail(PL_E_INVAL, "BGZF block without a valid BC (BSIZE) subfield");
    blk.resize(bsize);
    if (std::fread(blk.data() + 12 + xlen, 1, bsize - 12 - xlen, f) != bsize - 12 - xlen)
      return fail(PL_E_INVAL, "truncated BGZF block at byte %llu", (unsigned long long)c);
    Block b;
    b.cdata = cb.size() + 12 + xlen;
    b.clen = bsize - xlen - 20;
    b.crc = rd32(blk.data() + bsize - 8);
    b.isize = rd32(blk.data() + bsize - 4);
    b.uoff = utot;
    if (b.isize > 65536) return fail(PL_E_INVAL, "BGZF block ISIZE %u above 64 KiB", b.isize);
    utot += b.isize;
    cb.insert(cb.end(), blk.begin(), blk.end());
    blocks.push_back(b);
    coff.push_back(c);
    c += bsize;
  }
  return PL_OK;
}

pl_status open_indexed_impl(const char* path, const char* index, int nth, pl_bam* B) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(PL_E_INVAL, "cannot open %s", path);
  std::unique_ptr<FILE, int (*)(FILE*)> fh(f, std::fclose);
  B->path = path;
  B->nth = nth;
  B->indexed = true;
  // the header: blocks from the start until it parses
  std::vector<uint8_t> cb, ub;
  std::vector<Block> blocks;
  std::vector<uint64_t> coff;
  uint64_t at = 0;
  for (;;) {
    if (pl_status s = read_blocks(f, at, at, cb, blocks, coff)) return s;
    if (blocks.empty()) return fail(PL_E_INVAL, "not a BAM file (no complete header)");
    const size_t carry = ub.size();
    ub.resize(carry + blocks[0].isize);
    if (pl_status s = inflate_blocks(cb.data(), blocks, ub.data() + carry, 1)) return s;
    at = coff[0] + cb.size();
    const int64_t h = parse_header(ub.data(), ub.size(), B);
    if (h < 0) return fail(PL_E_INVAL, "not a BAM file (bad magic or header)");
    if (h > 0) break;
  }
  // the index (BamFile.scala:55-60: Pilon refuses a BAM without one)
  std::vector<std::string> cand;
  if (index && *index) cand.push_back(index);
  else {
    cand.push_back(std::string(path) + ".bai");
    const std::string p(path);
    if (p.size() > 4 && p.compare(p.size() - 4, 4, ".bam") == 0) cand.push_back(p.substr(0, p.size() - 4) + ".bai");
  }
  for (const std::string& c : cand) {
    FILE* t = std::fopen(c.c_str(), "rb");
    if (!t) continue;
    std::fclose(t);
    return read_bai(c, B->names.size(), B->bai);
  }
  return fail(PL_E_INVAL, "%s must be indexed BAM (no %s)", path, cand[0].c_str());
}

// Records of reference `ref` overlapping [start, stop] (1-based, closed;
// start <= 0 from the first base), through the index: the chunks of the
// bins reg2bins gives, after the linear index's lower bound, merged and read
// in file order; decoded like the whole-file path; kept if their span
// [pos, pos + reference length - 1] (one base for an unmapped or
// zero-length alignment) overlaps the interval.
pl_status query_impl(pl_bam* B, int32_t ref, int64_t start, int64_t stop) {
  B->query = RefReads();
  const BaiRef& I = B->bai[(size_t)ref];
  const int64_t beg = std::max<int64_t>(start - 1, 0), end = std::max<int64_t>(stop, beg);
  uint64_t min_off = 0;
  if (!I.lin.empty() && (size_t)(beg >> 14) < I.lin.size()) min_off = I.lin[(size_t)(beg >> 14)];
  std::vector<std::pair<uint64_t, uint64_t>> ch;
  for (uint32_t bin : reg2bins(beg, end)) {
    auto it = std::lower_bound(I.bins.begin(), I.bins.end(), bin,
                               [](const std::pair<uint32_t, std::vector<std::pair<uint64_t, uint64_t>>>& a,
                                  uint32_t b) { return a.first < b; });
    if (it == I.bins.end() || it->first != bin) continue;
    for (const auto& c : it->second)
      if (c.second > min_off) ch.push_back(c);
  }
  std::sort(ch.begin(), ch.end());
  std::vector<std::pair<uint64_t, uint64_t>> merged;
  for (const auto& c : ch) {
    if (!merged.empty() && c.first <= merged.back().second) merged.back().second = std::max(merged.back().second, c.second);
    else merged.push_back(c);
  }
  // decode the chunks into a scratch reader with the same dictionary
  pl_bam Q;
  Q.names = B->names;
  Q.lens = B->lens;
  Q.refs.resize(B->names.size());
  std::memset(&Q.st, 0, sizeof Q.st);
  FILE* f = std::fopen(B->path.c_str(), "rb");
  if (!f) return fail(PL_E_INVAL, "cannot open %s", B->path.c_str());
  std::unique_ptr<FILE, int (*)(FILE*)> fh(f, std::fclose);
  std::vector<uint8_t> cb, ub;
  std::vector<Block> blocks;
  std::vector<uint64_t> coff;
  for (const auto& c : merged) {
    const uint64_t c0 = c.first >> 16, c1 = c.second >> 16;
    if (pl_status s = read_blocks(f, c0, c1, cb, blocks, coff)) return s;
    if (blocks.empty() || coff[0] != c0) return fail(PL_E_INVAL, "index chunk outside the BAM");
    size_t utot = 0;
    for (const Block& b : blocks) utot += b.isize;
    ub.assign(utot, 0);
    if (pl_status s = inflate_blocks(cb.data(), blocks, ub.data(), B->nth)) return s;
    const size_t u0 = (size_t)(c.first & 0xffff);
    size_t u1 = utot;
    for (size_t k = 0; k < blocks.size(); k++)
      if (coff[k] == c1) { u1 = blocks[k].uoff + (size_t)(c.second & 0xffff); break; }
    if (u0 > u1 || u1 > utot) return fail(PL_E_INVAL, "index chunk outside the BAM");
    size_t used = u0;
    if (pl_status s = decode_records(&Q, ub.data(), u0, u1, B->nth, &used)) return s;
    if (used != u1) return fail(PL_E_INVAL, "index chunk ends inside a record");
  }
  // keep the reads of ref overlapping the interval
  const RefReads& R = Q.refs[(size_t)ref];
  RefReads& O = B->query;
  for (int64_t i = 0; i < R.n; i++) {
    int64_t rl = 0;
    for (int64_t k = R.cigar_off[i]; k < R.cigar_off[i + 1]; k++) {
      const uint32_t op = R.cigar[k] & 15u;
      if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rl += R.cigar[k] >> 4;
    }
    const int64_t ps = R.pos[i];
    const int64_t pe = ((R.flag[i] & 0x4) || rl == 0) ? ps : ps + rl - 1;
    if (ps > std::max<int64_t>(stop, 1) || pe < std::max<int64_t>(start, 1)) continue;
    const int64_t s0 = R.seq_off[i], L = R.seq_off[i + 1] - s0, oi = O.n;
    O.pos.push_back(R.pos[i]); O.tlen.push_back(R.tlen[i]); O.flag.push_back(R.flag[i]); O.mapq.push_back(R.mapq[i]);
    O.cigar.insert(O.cigar.end(), R.cigar.begin() + R.cigar_off[i], R.cigar.begin() + R.cigar_off[i + 1]);
    O.cigar_off.push_back((int64_t)O.cigar.size());
    const int64_t d0 = O.seq_off.back();
    O.seq_off.push_back(d0 + L);
    O.quals.insert(O.quals.end(), R.quals.begin() + s0, R.quals.begin() + s0 + L);
    O.bases2.resize((size_t)(d0 + L + 3) / 4, 0);
    if (R.flag[i] & PL_RF_EXOTIC) {
      const auto it = std::lower_bound(R.exo_read.begin(), R.exo_read.end(), i);
      const int64_t o = R.exo_off[(size_t)(it - R.exo_read.begin())];
      O.exo_read.push_back(oi);
      O.exo_off.push_back((int64_t)O.nib.size() * 2);
      O.nib.insert(O.nib.end(), R.nib.begin() + o / 2, R.nib.begin() + o / 2 + (L + 1) / 2);
    } else {
      for (int64_t k = 0; k < L; k++) {
        const int64_t a = s0 + k, b = d0 + k;
        O.bases2[(size_t)(b >> 2)] |= (uint8_t)(((R.bases2[(size_t)(a >> 2)] >> (2 * (a & 3))) & 3) << (2 * (b & 3)));
      }
    }
    O.n++;
  }
  if (O.nib.empty()) O.nib.push_back(0);
  B->st.records += Q.st.records;
  B->st.placed += O.n;
  B->st.inflate_s += Q.st.inflate_s;
  return PL_OK;
}

}  // namespace

extern "C" {

pl_status pl_bam_open_indexed(const char* path, const char* index, int32_t n_threads, pl_bam** out) {
  if (!path || !out) return fail(PL_E_INVAL, "null argument");
  *out = nullptr;
  int nth = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
  nth = std::max(1, std::min(nth, 64));
  std::unique_ptr<pl_bam> B(new (std::nothrow) pl_bam());
  if (!B) return fail(PL_E_OOM, "host allocation failed");
  std::memset(&B->st, 0, sizeof B->st);
  B->st.threads = nth;
  try {
    if (pl_status s = open_indexed_impl(path, index, nth, B.get())) return s;
  } catch (const std::bad_alloc&) {
    return fail(PL_E_OOM, "host allocation failed while opening %s", path);
  }
  *out = B.release();
  return PL_OK;
}

pl_status pl_bam_query(pl_bam* bam, int32_t ref, int64_t start, int64_t stop, pl_read_soa* out) {
  if (!bam || !out) return fail(PL_E_INVAL, "null argument");
  if (!bam->indexed) return fail(PL_E_INVAL, "the BAM was not opened with its index (pl_bam_open_indexed)");
  if (ref < 0 || ref >= (int32_t)bam->names.size()) return fail(PL_E_INVAL, "refID %d outside the dictionary", ref);
  try {
    if (pl_status s = query_impl(bam, ref, start, stop)) return s;
  } catch (const std::bad_alloc&) {
    return fail(PL_E_OOM, "host allocation failed in a query");
  }
  const RefReads& R = bam->query;
  std::memset(out, 0, sizeof *out);
  out->n_reads = R.n;
  out->pos = R.pos.data();
  out->tlen = R.tlen.data();
  out->flag = R.flag.data();
  out->mapq = R.mapq.data();
  out->cigar_off = R.cigar_off.data();
  out->cigar = R.cigar.data();
  out->seq_off = R.seq_off.data();
  out->bases2 = R.bases2.data();
  out->quals = R.quals.data();
  out->n_exotic = (int64_t)R.exo_read.size();
  out->exo_read = R.exo_read.data();
  out->exo_off = R.exo_off.data();
  out->nib = R.nib.data();
  return PL_OK;
}

pl_status pl_bam_open(const char* path, int32_t n_threads, pl_bam** out) {
  if (!path || !out) return fail(PL_E_INVAL, "null argument");
  *out = nullptr;
  int nth = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
  nth = std::max(1, std::min(nth, 64));
  std::unique_ptr<pl_bam> B(new (std::nothrow) pl_bam());
  if (!B) return fail(PL_E_OOM, "host allocation failed");
  std::memset(&B->st, 0, sizeof B->st);
  B->st.threads = nth;
  try {
    pl_status s = open_impl(path, nth, B.get());
    if (s) return s;
  } catch (const std::bad_alloc&) {
    return fail(PL_E_OOM, "host allocation failed while decoding %s", path);
  }
  *out = B.release();
  return PL_OK;
}

int32_t pl_bam_n_refs(const pl_bam* bam) { return bam ? (int32_t)bam->names.size() : 0; }

const char* pl_bam_ref_name(const pl_bam* bam, int32_t ref) {
  if (!bam || ref < 0 || ref >= (int32_t)bam->names.size()) return nullptr;
  return bam->names[(size_t)ref].c_str();
}

int64_t pl_bam_ref_len(const pl_bam* bam, int32_t ref) {
  if (!bam || ref < 0 || ref >= (int32_t)bam->lens.size()) return -1;
  return bam->lens[(size_t)ref];
}

pl_status pl_bam_ref_reads(const pl_bam* bam, int32_t ref, pl_read_soa* out) {
  if (!bam || !out) return fail(PL_E_INVAL, "null argument");
  if (ref < 0 || ref >= (int32_t)bam->refs.size()) return fail(PL_E_INVAL, "refID %d outside the dictionary", ref);
  if (bam->indexed) return fail(PL_E_INVAL, "an indexed reader reads records per region (pl_bam_query)");
  if (bam->streamed) return fail(PL_E_INVAL, "a stream hands over record bytes (pl_bam_stream_next)");
  const RefReads& R = bam->refs[(size_t)ref];
  std::memset(out, 0, sizeof *out);
  out->n_reads = R.n;
  out->pos = R.pos.data();
  out->tlen = R.tlen.data();
  out->flag = R.flag.data();
  out->mapq = R.mapq.data();
  out->cigar_off = R.cigar_off.data();
  out->cigar = R.cigar.data();
  out->seq_off = R.seq_off.data();
  out->bases2 = R.bases2.data();
  out->quals = R.quals.data();
  out->n_exotic = (int64_t)R.exo_read.size();
  out->exo_read = R.exo_read.data();
  out->exo_off = R.exo_off.data();
  out->nib = R.nib.data();
  return PL_OK;
}

pl_status pl_bam_stream_open(const char* path, int32_t n_threads, pl_bam** out) {
  if (!path || !out) return fail(PL_E_INVAL, "null argument");
  *out = nullptr;
  int nth = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
  nth = std::max(1, std::min(nth, 64));
  std::unique_ptr<pl_bam> B(new (std::nothrow) pl_bam());
  if (!B) return fail(PL_E_OOM, "host allocation failed");
  std::memset(&B->st, 0, sizeof B->st);
  B->st.threads = nth;
  B->nth = nth;
  B->streamed = true;
  try {
    if (pl_status s = map_file(path, B->map)) return s;
    B->st.compressed_bytes = (int64_t)B->map.n;
    // the inflated size: every block's ISIZE (the BSIZE chain)
    for (size_t q = 0; q + 18 <= B->map.n;) {
      const uint8_t* h = B->map.p + q;
      const size_t xlen = rd16(h + 10);
      size_t bsize = 0;
      for (size_t x = 12; x + 4 <= 12 + xlen && q + x + 6 <= B->map.n;) {
        const size_t slen = rd16(h + x + 2);
        if (h[x] == 66 && h[x + 1] == 67 && slen == 2) bsize = (size_t)rd16(h + x + 4) + 1;
        x += 4 + slen;
      }
      if (!bsize || q + bsize > B->map.n) break;      // next_blocks reports a bad block when it reaches it
      B->stream_bytes += rd32(h + bsize - 4);
      q += bsize;
    }
    // blocks until the header parses; what follows it is carried
    Buf<uint8_t> ub;
    for (;;) {
      std::vector<Block> blocks;
      size_t end = B->next_block;
      if (pl_status s = next_blocks(B->map, B->next_block, 1, blocks, &end, &B->last_at)) return s;   // one block
      if (blocks.empty()) return fail(PL_E_INVAL, "not a BAM file (no complete header)");
      B->last_end = end;
      size_t utot = 0;
      for (const Block& b : blocks) utot += b.isize;
      const size_t c0 = ub.size();
      ub.resize(c0 + utot);
      if (pl_status s = inflate_blocks(B->map.p, blocks, ub.data() + c0, nth)) return s;
      B->st.blocks += (int64_t)blocks.size();
      B->st.bam_bytes += (int64_t)utot;
      B->next_block = end;
      const int64_t h = parse_header(ub.data(), ub.size(), B.get());
      if (h < 0) return fail(PL_E_INVAL, "not a BAM file (bad magic or header)");
      if (h > 0) {
        B->carry.resize(ub.size() - (size_t)h);
        std::memcpy(B->carry.data(), ub.data() + h, ub.size() - (size_t)h);
        break;
      }
    }
  } catch (const std::bad_alloc&) {
    return fail(PL_E_OOM, "host allocation failed while opening %s", path);
  }
  *out = B.release();
  return PL_OK;
}

pl_status pl_bam_stream_next(pl_bam* bam, uint8_t* dst, int64_t cap, int64_t* len, int64_t* rec_off, int64_t rec_cap,
                             int64_t* n_rec) {
  if (!bam || !dst || !len || !rec_off || !n_rec) return fail(PL_E_INVAL, "null argument");
  if (!bam->streamed) return fail(PL_E_INVAL, "not a stream (pl_bam_stream_open)");
  try {
    return stream_next_impl(bam, dst, cap, len, rec_off, rec_cap, n_rec);
  } catch (const std::bad_alloc&) {
    return fail(PL_E_OOM, "host allocation failed");
  }
}

int64_t pl_bam_stream_bytes(const pl_bam* bam) { return bam && bam->streamed ? bam->stream_bytes : -1; }

void pl_bam_get_stats(const pl_bam* bam, pl_bam_stats* out) {
  if (bam && out) *out = bam->st;
}

void pl_bam_close(pl_bam* bam) { delete bam; }

const char* pl_bam_last_error(void) { return g_err.c_str(); }

}  // extern "C"
