// pl_core.h — the per-read and per-column logic of the pileup path, written
// once as __host__ __device__ code.  The gfx950 kernels call it with LDS and
// global-atomic sinks; a host build of the same functions (core_host.cpp)
// exists only so CPU tests can check the window decomposition.
//
// Reference: broadinstitute/pilon v1.24, src/main/scala/org/broadinstitute/pilon/
#pragma once
#include <stdint.h>
#include <limits.h>
#include <math.h>
#include "pl_common.h"

#ifndef PL_HD
#define PL_HD __host__ __device__ __forceinline__
#endif

// counter bump that is an atomic on the device and a plain add on the host
PL_HD uint32_t pl_bump(uint32_t* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, 1u);
#else
  return (*p)++;
#endif
}

// ------------------------------------------------------------ Java ints
PL_HD int32_t j_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
PL_HD int32_t j_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
PL_HD int32_t j_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
PL_HD int32_t j_div(int32_t a, int32_t b) { return (a == INT_MIN && b == -1) ? a : a / b; }
// Utils.roundDiv / pct (Utils.scala:23-26)
PL_HD int32_t round_div_i(int32_t n, int32_t d) { return d > 0 ? j_div(j_add(n, d / 2), d) : 0; }
PL_HD int64_t round_div_l(int64_t n, int64_t d) {
  if (d <= 0) return 0;
  const int64_t x = n + d / 2;
  // a non-negative quotient of 32-bit operands: one 32-bit division
  if ((uint64_t)x <= 0xffffffffull && (uint64_t)d <= 0xffffffffull) return (int64_t)((uint32_t)x / (uint32_t)d);
  return x / d;
}
PL_HD int32_t pct_i(int32_t n, int32_t d) { return round_div_i(j_mul(100, n), d); }

// ------------------------------------------------------------ bases
// getReadBases decodes BAM nibbles through "=ACMGRSVTWYHKDBN"
PL_HD uint8_t nib_char(int c) {
  switch (c & 15) {
    case 0: return '='; case 1: return 'A'; case 2: return 'C'; case 3: return 'M';
    case 4: return 'G'; case 5: return 'R'; case 6: return 'S'; case 7: return 'V';
    case 8: return 'T'; case 9: return 'W'; case 10: return 'Y'; case 11: return 'H';
    case 12: return 'K'; case 13: return 'D'; case 14: return 'B'; default: return 'N';
  }
}
PL_HD uint8_t acgt_char(int code) { return (uint8_t)((0x54474341u >> (8 * (code & 3))) & 0xff); }
// PileUp.baseIndex (PileUp.scala:46-52)
PL_HD int base_index(uint8_t c) {
  return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : -1;
}
PL_HD uint8_t to_upper(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }

struct ReadAcc {           // bases/quals of one read
  const uint8_t* bases2;
  const uint8_t* nib;
  const int8_t* quals;
  int64_t seq0;            // first base in the 2-bit/qual streams
  int64_t nib0;            // first nibble (exotic reads)
  bool exotic, noqual;
  int32_t default_qual;
  PL_HD uint8_t ch(int32_t o) const {
    if (!exotic) {
      int64_t q = seq0 + o;
      return acgt_char((bases2[q >> 2] >> (2 * (q & 3))) & 3);
    }
    int64_t q = nib0 + o;
    uint8_t b = nib[q >> 1];
    return nib_char((q & 1) ? (b & 15) : (b >> 4));
  }
  PL_HD int bi(int32_t o) const {
    if (!exotic) {
      int64_t q = seq0 + o;
      return (bases2[q >> 2] >> (2 * (q & 3))) & 3;
    }
    return base_index(ch(o));
  }
  PL_HD int32_t qual(int32_t o) const { return noqual ? default_qual : (int32_t)quals[seq0 + o]; }
};

PL_HD int64_t exotic_nib0(const RSDev& R, int64_t r) {
  int64_t lo = 0, hi = R.n_exotic;
  while (lo < hi) { int64_t m = (lo + hi) >> 1; if (R.exo_read[m] < r) lo = m + 1; else hi = m; }
  return (lo < R.n_exotic && R.exo_read[lo] == r) ? R.exo_off[lo] : 0;
}

PL_HD ReadAcc read_acc(const RSDev& R, int64_t r, uint32_t info, int32_t default_qual) {
  ReadAcc a;
  a.bases2 = R.bases2; a.nib = R.nib; a.quals = R.quals;
  a.seq0 = R.seq_off[r];
  a.exotic = (info & RI_EXOTIC) != 0;
  a.noqual = (info & RI_NOQUAL) != 0;
  a.nib0 = a.exotic ? exotic_nib0(R, r) : 0;
  a.default_qual = default_qual;
  return a;
}

PL_HD bool consumes_read(int op) { return op == 0 || op == 1 || op == 4 || op == 7 || op == 8; }
PL_HD bool consumes_ref(int op) { return op == 0 || op == 2 || op == 3 || op == 7 || op == 8; }

// ------------------------------------------------------------ normalisation
// Insertion left-normalisation (PileUpRegion.scala:153-157): slide while the
// base before the insertion equals its last (rotated) base.  Returns the final
// locus; *rot counts right rotations (mod len).
PL_HD int32_t slide_insertion(const uint8_t* ref, const ReadAcc& a, int32_t readOffset, int32_t len,
                              int32_t locus, int32_t* rot) {
  int32_t iloc = locus, k = 0;   // rotated[j] = orig[(j - k) mod len]
  while (iloc > 1) {
    int32_t lastIdx = (len - 1 - k) % len;
    if (lastIdx < 0) lastIdx += len;
    if (ref[iloc - 2] != a.ch(readOffset + lastIdx)) break;
    iloc -= 1;
    k = (k + 1) % len;
  }
  *rot = k;
  return iloc;
}

// FNV-1a over the rotated insertion string; flags bit1 if it contains 'N'.
PL_HD uint32_t ins_hash(const ReadAcc& a, int32_t readOffset, int32_t len, int32_t rot, uint32_t* hasN) {
  uint32_t h = 2166136261u;
  *hasN = 0;
  for (int32_t j = 0; j < len; j++) {
    int32_t oi = (j - rot) % len;
    if (oi < 0) oi += len;
    uint8_t c = a.ch(readOffset + oi);
    if (c == 'N') *hasN = 1;
    h = (h ^ c) * 16777619u;
  }
  return h;
}
PL_HD uint32_t del_hash(const uint8_t* ref, int64_t off, int32_t len, uint32_t* hasN) {
  uint32_t h = 2166136261u;
  *hasN = 0;
  for (int32_t j = 0; j < len; j++) {
    uint8_t c = ref[off + j];
    if (c == 'N') *hasN = 1;
    h = (h ^ c) * 16777619u;
  }
  return h;
}

// ------------------------------------------------------------ per-read prep
// Per-read facts every region needs, computed once per resident read set:
// alignment end (htsjdk getAlignmentEnd), validity and adjMq
// (PileUpRegion.scala:107,139-141), the span of positions the read can touch
// (M bases, soft clips, slid indels) and its physical-coverage fragment
// (physCovIncr, PileUpRegion.scala:62-73).
struct PrepOut { RMeta m; int32_t events; };

PL_HD PrepOut prep_read(const RSDev& R, int64_t r, const KConf& cf) {
  PrepOut o;
  const uint16_t flag = R.flag[r];
  const int32_t mq = R.mapq[r];
  const int32_t pos = R.pos[r];
  const int32_t tlen = R.tlen[r];
  const int64_t c0 = R.cigar_off[r], c1 = R.cigar_off[r + 1];
  const int32_t L = (int32_t)(R.seq_off[r + 1] - R.seq_off[r]);
  uint32_t info = 0;
  int32_t refLen = 0, readLen = 0, clipped = 0, nP = 0;
  bool bad = false, hasS = false;
  for (int64_t k = c0; k < c1; k++) {
    uint32_t w = R.cigar[k];
    int op = (int)(w & 15);
    int32_t ln = (int32_t)(w >> 4);
    if (op > 8) bad = true;
    if (consumes_ref(op)) refLen += ln;
    if (consumes_read(op)) readLen += ln;
    if (op == 4) { clipped = j_add(clipped, ln); hasS = true; }
    if (op == 6) nP++;
  }
  if (L > 0 && readLen != L) bad = true;
  const int32_t aend = (flag & 0x4) ? 0 : pos + refLen - 1;
  const bool keep = (cf.nonpf || !(flag & 0x200)) && (cf.duplicates || !(flag & 0x400)) && !(flag & 0x100);
  const bool paired = (flag & 0x1) != 0;
  const bool valid = (mq >= cf.min_mq) && (!paired || (((flag & 0x2) != 0) && ((flag & PL_RF_MATE_SAME_REF) != 0)));
  const int32_t adjMq = round_div_i(j_mul(mq, j_sub(L, clipped)), L);
  info |= (uint32_t)adjMq & RI_ADJMQ_MASK;
  if (valid) info |= RI_VALID;
  if (keep) info |= RI_KEEP;
  if (flag & PL_RF_EXOTIC) info |= RI_EXOTIC;
  if (flag & PL_RF_NO_QUAL) info |= RI_NOQUAL;
  if (bad) info |= RI_BAD;
  if (L == 0 && hasS) info |= RI_EDGE_CLIP;
  info |= (uint32_t)(nP > 255 ? 255 : nP) << RI_UNK_SHIFT;
  // a long read is always walked (its bases and indels depend on the
  // reference around them): never simple, and too many segments
  const uint32_t lr = R.lrt ? (uint32_t)R.lrt[r] : 0u;
  if (lr) info |= (lr << RI_LR_SHIFT) | ((uint32_t)RI_SEG_MAX << RI_SEG_SHIFT);
  const uint32_t op0 = c1 > c0 ? R.cigar[c0] : 0u;
  if (c1 - c0 == 1 && L > 0 && !bad && !(flag & PL_RF_EXOTIC) && !lr) {
    const int op = (int)(op0 & 15);
    if (op == 0 || op == 7 || op == 8) info |= RI_SIMPLE;
  }

  int32_t lo = INT_MAX, hi = INT_MIN, events = 0;
  if (!bad) {
    ReadAcc a = read_acc(R, r, info, cf.default_qual);
    const int32_t flank = cf.flank;
    int32_t readOffset = 0, refOffset = 0;
    for (int64_t k = c0; k < c1; k++) {
      uint32_t w = R.cigar[k];
      int op = (int)(w & 15);
      int32_t len = (int32_t)(w >> 4);
      int32_t locus = pos + refOffset;
      bool tr = readOffset >= flank && j_sub(L, flank) > readOffset;
      if ((op == 0 || op == 7 || op == 8) && len > 0) {
        lo = locus < lo ? locus : lo;
        hi = locus + len - 1 > hi ? locus + len - 1 : hi;
      } else if (op == 4 && len > 0) {
        int32_t cs = readOffset == 0 ? locus - len : locus;
        lo = cs < lo ? cs : lo;
        hi = cs + len - 1 > hi ? cs + len - 1 : hi;
      } else if (op == 1 && len > 0 && valid && tr && locus <= R.ref_len) {
        int32_t rot;
        int32_t il = slide_insertion(R.ref, a, readOffset, len, locus, &rot);
        lo = il < lo ? il : lo;
        hi = locus > hi ? locus : hi;
        events++;
      } else if (op == 2 && len > 0 && valid && tr && (int64_t)locus + len - 1 <= R.ref_len) {
        int32_t dl = locus, rl = readOffset;
        while (dl > 1 && rl > 0 && R.ref[dl - 2] == R.ref[dl + len - 2]) { dl--; rl--; }
        lo = dl < lo ? dl : lo;
        hi = locus + len - 1 > hi ? locus + len - 1 : hi;
        events++;
      }
      if (consumes_read(op)) readOffset += len;
      if (consumes_ref(op)) refOffset += len;
    }
  }
  if (events) info |= RI_EVENTS;
  // physCovIncr fragment (only valid reads, and not the right mate of a pair)
  int32_t flo = 0, fhi = 0;
  if (valid && keep && !(paired && tlen <= 0)) {
    if (!paired) { flo = pos < aend ? pos : aend; fhi = pos > aend ? pos : aend; }
    else { flo = pos; fhi = j_add(pos, tlen); }
  }
  o.m.pos = pos; o.m.info = info; o.m.fp_lo = lo; o.m.fp_hi = hi;
  o.m.fr_lo = flo; o.m.fr_hi = fhi; o.m.aend = aend; o.m.op0 = op0;
  o.events = events;
  return o;
}

// ------------------------------------------------------------ the read walk
// addRead's homoRun(i0) >= 4 (PileUpRegion.scala:120-126): the contig byte
// at 0-based index i0 (0 <= i0 < ref_len) and the three after it are equal.
PL_HD bool homo_run4(const RSDev& R, int64_t i0) {
  if (i0 + 3 >= R.ref_len) return false;
  const uint8_t b = R.ref[i0];
  return R.ref[i0 + 1] == b && R.ref[i0 + 2] == b && R.ref[i0 + 3] == b;
}

// addRead's nanoporeExclude(i0) (PileUpRegion.scala:128-134): i0 is a region
// index (locus - start) but indexes the contig's bytes; the region test is
// on locus(i0 -+ 2).
PL_HD bool nanopore_exclude(const RSDev& R, const JobDev& J, int64_t i0) {
  if (i0 < 2 || (int64_t)J.start + i0 + 2 > J.stop) return false;
  return R.ref[i0 - 2] == 'C' && R.ref[i0 - 1] == 'C' && R.ref[i0 + 1] == 'G' && R.ref[i0 + 2] == 'G';
}

// Insertion event of one I op (PileUpRegion.scala:150-162): gate, left
// normalisation, then addInsertion at the final locus if it lies in the window.
// A long read's insertion (lr > 0) is dropped where homoRun(iloc) >= 4, and
// homoRun reading past the contig's end throws (:160).
template <class Sink>
PL_HD void ins_event(const RSDev& R, const JobDev& J, const ReadAcc& a, int64_t r, int32_t locus, int32_t readOffset,
                     int32_t len, int32_t mq1, int32_t L0, int32_t L1, Sink& s, uint32_t lr = 0) {
  int32_t rot;
  const int32_t il = slide_insertion(R.ref, a, readOffset, len, locus, &rot);
  if (lr) {
    if ((int64_t)il >= R.ref_len) { s.edge(r); return; }
    if (homo_run4(R, il)) return;
  }
  if (il < J.start) { s.edge(r); return; }
  if (il >= L0 && il < L1) {
    uint32_t hasN;
    const uint32_t h = ins_hash(a, readOffset, len, rot, &hasN);
    IndelRec e;
    e.key = (uint32_t)(((uint64_t)(J.out + (il - J.start)) << 1) | 0u);
    e.len = len; e.qual = (int16_t)a.qual(readOffset); e.mq1 = (uint16_t)mq1;
    e.hash = h; e.rot = rot; e.flags = (a.exotic ? 1u : 0u) | (hasN ? 2u : 0u);
    e.src = (a.exotic ? a.nib0 : a.seq0) + readOffset;
    s.ins(il - L0, a.qual(readOffset), mq1, e);
  }
}

// Deletion event of one D op (PileUpRegion.scala:163-183): left normalisation
// with the double-counting add of each slid-over base (remove() is a no-op for
// valid reads), then addDeletion at the final locus.  mq1 is adjMq + 1 for
// the slid-over bases, imq1 indelMq + 1 for the deletion; a long read's
// deletion is dropped where homoRun(index(dloc)) >= 4 or, for nanopore,
// nanoporeExclude(index(dloc)) holds (:180-182).
template <class Sink>
PL_HD void del_event(const RSDev& R, const JobDev& J, const KConf& cf, const ReadAcc& a, int64_t r, int32_t locus,
                     int32_t readOffset, int32_t len, int32_t L, int32_t mq1, int32_t L0, int32_t L1, Sink& s,
                     uint32_t lr = 0, int32_t imq1 = 0) {
  if (!lr) imq1 = mq1;
  int32_t dl = locus, rl = readOffset;
  const int32_t flank = cf.flank;
  while (dl > 1 && rl > 0 && R.ref[dl - 2] == R.ref[dl + len - 2]) {
    dl--; rl--;
    const bool trl = rl >= flank && j_sub(L, flank) > rl;
    if (trl && dl >= J.start && dl <= J.stop && dl + len >= J.start && dl + len <= J.stop) {
      const int32_t t = dl + len;
      if (t >= L0 && t < L1) {
        const int b = a.bi(rl);
        const int32_t q = a.qual(rl);
        if (b >= 0 && q >= cf.min_qual) s.base(t - L0, b, q, mq1);
        s.count_base();
      }
    }
  }
  if (dl < J.start) { s.edge(r); return; }
  if (lr) {
    const int64_t i0 = (int64_t)dl - J.start;
    if (homo_run4(R, i0) || (lr == PL_LR_NANOPORE && nanopore_exclude(R, J, i0))) return;
  }
  if (dl >= L0 && dl < L1) {
    uint32_t hasN;
    const uint32_t h = del_hash(R.ref, (int64_t)dl - 1, len, &hasN);
    IndelRec e;
    e.key = (uint32_t)(((uint64_t)(J.out + (dl - J.start)) << 1) | 1u);
    e.len = len; e.qual = (int16_t)a.qual(readOffset); e.mq1 = (uint16_t)imq1;
    e.hash = h; e.rot = 0; e.flags = hasN ? 2u : 0u;
    e.src = (int64_t)dl - 1;
    s.del(dl - L0, a.qual(readOffset), imq1, e);
  }
}

// Is the read one BamFile.process hands to addRead for this region?
PL_HD bool read_in_query(const RMeta& m, const JobDev& J) {
  if (!(m.info & RI_KEEP) || (m.info & RI_BAD)) return false;
  const int32_t spanEnd = m.aend != 0 ? m.aend : m.pos;
  return m.pos <= J.qe && spanEnd >= J.qs;                 // queryOverlapping (BamFile.scala:118-119)
}

// physCovIncr restricted to window [L0, L1): the region's prefix sum of these
// differences is #{fragments: max(start, region.start) <= i < end}.
template <class Sink>
PL_HD void frag_event(const RMeta& m, const JobDev& J, int32_t L0, int32_t L1, Sink& s) {
  if (m.fr_lo < m.fr_hi) {
    int32_t a = m.fr_lo > J.start ? m.fr_lo : J.start;
    a = a > L0 ? a : L0;
    if (a < m.fr_hi && a < L1) {
      const int32_t size = j_sub(m.fr_hi, m.fr_lo);
      s.frag(a - L0, 1, size);
      if (m.fr_hi < L1) s.frag(m.fr_hi - L0, -1, j_sub(0, size));
    }
  }
}

// What frag_event needs of a read that reaches a window only through its
// fragment (computed once per read set, after the outlier decision).
PL_HD FragTail frag_tail(const RMeta& m) {
  FragTail t;
  t.hi = INT32_MIN;
  t.size = 0;
  if ((m.info & RI_KEEP) && !(m.info & (RI_BAD | RI_OUTLIER)) && m.fr_lo < m.fr_hi) {
    t.hi = m.fr_hi;
    t.size = j_sub(m.fr_hi, m.fr_lo);
  }
  return t;
}

// End of the window's fragment-only reads: [lo, end) holds the reads with
// pos < L0 - back_fp.  Their pileup footprint ends before L0 (fp_hi <= pos +
// back_fp), and their fragment starts at or before pos (fr_lo <= pos), so
// frag_event reduces to +1/+size at L0 and -1/-size at the fragment end,
// which frag_tail holds.  read_in_query reduces to the kept/valid flags when
// every such read has pos - 1 >= qs (spanEnd >= pos - 1) and pos < L0 <= qe;
// otherwise the window takes every read through the full path.
PL_HD int64_t fragment_only_end(const int32_t* pos, int64_t lo, int64_t hi, const JobDev& J, int32_t L0,
                                int32_t back, int32_t back_fp) {
  if ((int64_t)L0 - back - 1 < (int64_t)J.qs || back_fp >= back) return lo;
  const int64_t x = (int64_t)L0 - back_fp;
  while (lo < hi) {
    const int64_t m = (lo + hi) >> 1;
    if ((int64_t)pos[m] < x) lo = m + 1; else hi = m;
  }
  return lo;
}

// frag_event of a fragment-only read from its tail: returns 1 when it adds
// +1/+size at window position 0 (the caller sums those), and applies the
// -1/-size at the fragment end when that falls inside the window.
template <class Sink>
PL_HD int frag_tail_event(const FragTail& t, int32_t L0, int32_t L1, Sink& s) {
  if (t.hi <= L0) return 0;
  if (t.hi < L1) s.frag(t.hi - L0, -1, j_sub(0, t.size));
  return 1;
}

// PileUpRegion.addRead restricted to the positions of one window [L0, L1)
// (a slice of region [J.start, J.stop]).  Lanes stride over the bases of an
// op; lane 0 handles the per-op events.  The Sink receives integer
// increments only, so the result does not depend on read order.
template <class Sink>
PL_HD void walk_read(const RSDev& R, const JobDev& J, const KConf& cf, int64_t r, int32_t L0, int32_t L1,
                     int lane, int nlanes, Sink& s, bool do_frag = true) {
  const RMeta m = R.meta[r];
  if (!read_in_query(m, J)) return;
  if (do_frag && lane == 0) frag_event(m, J, L0, L1, s);
  if (m.fp_hi < L0 || m.fp_lo >= L1) return;
  const uint32_t info = m.info;
  const int32_t pos = m.pos;
  const bool valid = (info & RI_VALID) != 0;
  const int32_t adjMq = (int32_t)(info & RI_ADJMQ_MASK);
  const int32_t mq1 = adjMq + 1;
  const uint32_t lr = (info & RI_LR_MASK) >> RI_LR_SHIFT;
  const int32_t imq1 = (lr && adjMq > 8 ? 8 : adjMq) + 1;      // indelMq (PileUpRegion.scala:142)
  const int32_t L = (int32_t)(R.seq_off[r + 1] - R.seq_off[r]);
  const int32_t flank = cf.flank;
  const ReadAcc a = read_acc(R, r, info, cf.default_qual);
  const int64_t c0 = R.cigar_off[r], c1 = R.cigar_off[r + 1];
  int32_t readOffset = 0, refOffset = 0;
  for (int64_t k = c0; k < c1; k++) {
    const uint32_t w = R.cigar[k];
    const int op = (int)(w & 15);
    const int32_t len = (int32_t)(w >> 4);
    const int32_t locus = pos + refOffset;
    if (op == 0 || op == 7 || op == 8) {                       // M/=/X (:184-193)
      int32_t ab = (locus > L0 ? locus : L0), ae = (locus + len < L1 ? locus + len : L1);
      if (ae > ab && lane == 0) s.aligned(ae - ab);
      int32_t i0 = L0 - locus;
      if (i0 < 0) i0 = 0;
      if (flank - readOffset > i0) i0 = flank - readOffset;
      int32_t i1 = len;
      if (L1 - locus < i1) i1 = L1 - locus;
      if (L - flank - readOffset < i1) i1 = L - flank - readOffset;
      for (int32_t i = i0 + lane; i < i1; i += nlanes) {
        const int32_t o = readOffset + i;
        const int32_t p = locus + i - L0;
        if (valid) {
          const int b = a.bi(o);
          // a nanopore base inside CC_GG piles with quality 0 (:190)
          const int32_t q = lr == PL_LR_NANOPORE && nanopore_exclude(R, J, (int64_t)locus + i - J.start) ? 0 : a.qual(o);
          if (b >= 0 && q >= cf.min_qual) s.base(p, b, q, mq1);   // PileUp.add (PileUp.scala:75-84)
          s.count_base();                                           // region baseCount (:42)
        } else {
          s.bad(p, 1);                                              // badPair (:44)
        }
      }
    } else if (op == 4) {                                      // S (:194-206)
      const int32_t cs = readOffset == 0 ? locus - len : locus;
      const int32_t ce = cs + len - 1;
      if (lane == 0) {
        if (cs >= L0 && cs < L1) s.clip(cs - L0);
        if (ce >= L0 && ce < L1) s.clip(ce - L0);
      }
      int32_t i0 = L0 - cs, i1 = L1 - cs;
      if (i0 < 0) i0 = 0;
      if (i1 > len) i1 = len;
      for (int32_t i = i0 + lane; i < i1; i += nlanes) s.bad(cs + i - L0, 1);
    } else if (op == 1) {                                      // I (:150-162)
      const bool tr = readOffset >= flank && j_sub(L, flank) > readOffset;
      if (lane == 0 && valid && tr && locus >= J.start && locus <= J.stop)
        ins_event(R, J, a, r, locus, readOffset, len, imq1, L0, L1, s, lr);
    } else if (op == 2) {                                      // D (:163-183)
      const bool tr = readOffset >= flank && j_sub(L, flank) > readOffset;
      if (lane == 0 && valid && tr && locus >= J.start && locus <= J.stop &&
          locus + len - 1 >= J.start && locus + len - 1 <= J.stop)
        del_event(R, J, cf, a, r, locus, readOffset, len, L, mq1, L0, L1, s, lr, imq1);
    }
    if (consumes_read(op)) readOffset += len;
    if (consumes_ref(op)) refOffset += len;
  }
}

// ------------------------------------------------------------ flattened runs
// A run of consecutive bases of one read inside the window for the flattened
// base loop: M/=/X bases (trusted, valid reads) or positions that only take a
// badPair increment (clipped bases, bases of invalid reads).
#define EF_VALID  (1u << 25)
#define EF_NOQUAL (1u << 26)
#define EF_BAD    (1u << 27)
struct Run {
  int32_t p0;        // window index of the first base
  int32_t n;
  int64_t q;         // base-stream index of the first base
  uint32_t fl;       // mq1 << 16 | EF_* bits
};

// The run of a single-op read (RI_SIMPLE): its trusted bases in the window.
PL_HD int simple_run(const RMeta& m, int64_t seq0, const KConf& cf, int32_t L0, int32_t L1, Run* out,
                     int64_t* aligned, int64_t* bcount) {
  const int32_t pos = m.pos;
  const int32_t len = (int32_t)(m.op0 >> 4);   // == L for a simple read
  const int32_t ab = pos > L0 ? pos : L0, ae = pos + len < L1 ? pos + len : L1;
  if (ae > ab) *aligned += ae - ab;
  int32_t o0 = L0 - pos, o1 = L1 - pos;
  if (o0 < cf.flank) o0 = cf.flank;
  if (o1 > len - cf.flank) o1 = len - cf.flank;
  if (o1 <= o0) return 0;
  const bool valid = (m.info & RI_VALID) != 0;
  out->p0 = pos + o0 - L0;
  out->n = o1 - o0;
  out->q = seq0 + o0;
  out->fl = valid ? ((((m.info & RI_ADJMQ_MASK) + 1u) << 16) | EF_VALID | ((m.info & RI_NOQUAL) ? EF_NOQUAL : 0u))
                  : EF_BAD;
  if (valid) *bcount += o1 - o0;
  return 1;
}


// ------------------------------------------------------------ segment table
// The region-independent part of a multi-op read, computed once per read set:
// its M/S ops, and its insertion/deletion events with Pilon's left
// normalisation already applied (final locus, rotation, string hash) and,
// for deletions, each slid-over base that addRead adds len bases later.
// Region gates (inRegion) are applied at window time.
enum { SEG_M = 0, SEG_S = 1, SEG_I = 2, SEG_D = 3, SEG_A = 4 };
struct __attribute__((aligned(16))) Seg {
  int32_t kq;        // kind | quals(readOffset or rl) << 8
  int32_t locus;     // M: first locus; S: clipStart; I/D: locus before normalisation; A: slid dloc
  int32_t roff;      // read offset (A: rl)
  int32_t len;       // op length (A: deletion length)
  int32_t fin;       // I: final iloc; D: final dloc; A: base index of read base rl (-1 if not ACGT)
  uint32_t hash;     // I/D: FNV-1a of the final string
  int32_t rotn;      // I: rotations | hasN << 30; D: hasN << 30
  int32_t pad;
};
PL_HD int seg_kind(const Seg& g) { return g.kq & 0xff; }
PL_HD int32_t seg_qual(const Seg& g) { return g.kq >> 8; }

// Writes the segments of read r into out (if non-null); returns their count.
PL_HD int build_segs(const RSDev& R, int64_t r, uint32_t info, const KConf& cf, Seg* out) {
  const int64_t c0 = R.cigar_off[r], c1 = R.cigar_off[r + 1];
  const int32_t L = (int32_t)(R.seq_off[r + 1] - R.seq_off[r]);
  const int32_t pos = R.pos[r];
  const bool valid = (info & RI_VALID) != 0;
  const int32_t flank = cf.flank;
  const ReadAcc a = read_acc(R, r, info, cf.default_qual);
  int n = 0;
  int32_t readOffset = 0, refOffset = 0;
  for (int64_t k = c0; k < c1; k++) {
    const uint32_t w = R.cigar[k];
    const int op = (int)(w & 15);
    const int32_t len = (int32_t)(w >> 4);
    const int32_t locus = pos + refOffset;
    const bool tr = readOffset >= flank && j_sub(L, flank) > readOffset;
    if ((op == 0 || op == 7 || op == 8) && len > 0) {
      if (out) { Seg& g = out[n]; g.kq = SEG_M; g.locus = locus; g.roff = readOffset; g.len = len; g.fin = 0; g.hash = 0; g.rotn = 0; g.pad = 0; }
      n++;
    } else if (op == 4 && len > 0) {
      if (out) {
        Seg& g = out[n];
        g.kq = SEG_S; g.locus = readOffset == 0 ? locus - len : locus; g.roff = readOffset; g.len = len;
        g.fin = 0; g.hash = 0; g.rotn = 0; g.pad = 0;
      }
      n++;
    } else if (op == 1 && len > 0 && valid && tr && locus <= R.ref_len) {
      if (out) {
        int32_t rot;
        const int32_t il = slide_insertion(R.ref, a, readOffset, len, locus, &rot);
        uint32_t hasN;
        const uint32_t h = ins_hash(a, readOffset, len, rot, &hasN);
        Seg& g = out[n];
        g.kq = SEG_I | (a.qual(readOffset) << 8); g.locus = locus; g.roff = readOffset; g.len = len;
        g.fin = il; g.hash = h; g.rotn = rot | (int32_t)(hasN << 30); g.pad = 0;
      }
      n++;
    } else if (op == 2 && len > 0 && valid && tr && (int64_t)locus + len - 1 <= R.ref_len) {
      int32_t dl = locus, rl = readOffset;
      const int d_at = n++;
      while (dl > 1 && rl > 0 && R.ref[dl - 2] == R.ref[dl + len - 2]) {
        dl--; rl--;
        if (rl >= flank && j_sub(L, flank) > rl) {
          if (out) {
            Seg& g = out[n];
            g.kq = SEG_A | (a.qual(rl) << 8); g.locus = dl; g.roff = rl; g.len = len;
            g.fin = a.bi(rl); g.hash = 0; g.rotn = 0; g.pad = locus;   // pad: the deletion's own locus (its gate)
          }
          n++;
        }
      }
      if (out) {
        uint32_t hasN;
        const uint32_t h = del_hash(R.ref, (int64_t)dl - 1, len, &hasN);
        Seg& g = out[d_at];
        g.kq = SEG_D | (a.qual(readOffset) << 8); g.locus = locus; g.roff = readOffset; g.len = len;
        g.fin = dl; g.hash = h; g.rotn = (int32_t)(hasN << 30); g.pad = 0;
      }
    }
    if (consumes_read(op)) readOffset += len;
    if (consumes_ref(op)) refOffset += len;
  }
  return n;
}

// One segment of a multi-op read at window time, applied on its own: every
// segment carries what its gates need (a slide add carries its deletion's
// locus), so the segments of a read can be applied in any order or in
// parallel.  Clips, indel events and slide adds go to the sink with the
// region gates of PileUpRegion.addRead; an M segment with trusted in-window
// bases, or an in-window clip, returns 1 and its run in ru (clip runs and
// runs of invalid reads carry EF_BAD: badPair only).
template <class Sink>
PL_HD int seg_apply(const JobDev& J, const KConf& cf, int64_t r, uint32_t info, int64_t seq0, int32_t L,
                    int32_t L0, int32_t L1, const Seg& g, Run& ru, int64_t* aligned, int64_t* bcount, Sink& s) {
  const bool valid = (info & RI_VALID) != 0;
  const int32_t mq1 = (int32_t)(info & RI_ADJMQ_MASK) + 1;
  const int32_t flank = cf.flank;
  const int kd = seg_kind(g);
  if (kd == SEG_M) {                                       // (:184-193)
    const int32_t ab = g.locus > L0 ? g.locus : L0, ae = g.locus + g.len < L1 ? g.locus + g.len : L1;
    if (ae > ab) *aligned += ae - ab;
    int32_t i0 = L0 - g.locus, i1 = g.len;
    if (i0 < 0) i0 = 0;
    if (flank - g.roff > i0) i0 = flank - g.roff;
    if (L1 - g.locus < i1) i1 = L1 - g.locus;
    if (L - flank - g.roff < i1) i1 = L - flank - g.roff;
    if (i1 <= i0) return 0;
    ru.p0 = g.locus + i0 - L0; ru.n = i1 - i0; ru.q = seq0 + g.roff + i0;
    ru.fl = valid ? (((uint32_t)mq1 << 16) | EF_VALID | ((info & RI_NOQUAL) ? EF_NOQUAL : 0u)) : EF_BAD;
    if (valid) *bcount += i1 - i0;
    return 1;
  }
  if (kd == SEG_S) {                                       // (:194-206)
    const int32_t cs = g.locus, ce = cs + g.len - 1;
    if (cs >= L0 && cs < L1) s.clip(cs - L0);
    if (ce >= L0 && ce < L1) s.clip(ce - L0);
    int32_t i0 = L0 - cs, i1 = L1 - cs;
    if (i0 < 0) i0 = 0;
    if (i1 > g.len) i1 = g.len;
    if (i1 <= i0) return 0;
    ru.p0 = cs + i0 - L0; ru.n = i1 - i0; ru.q = seq0; ru.fl = EF_BAD;
    return 1;
  }
  if (kd == SEG_I) {                                       // (:150-162)
    if (g.locus >= J.start && g.locus <= J.stop) {
      const int32_t il = g.fin;
      if (il < J.start) s.edge(r);
      else if (il >= L0 && il < L1) {
        IndelRec e;
        e.key = (uint32_t)(((uint64_t)(J.out + (il - J.start)) << 1) | 0u);
        e.len = g.len; e.qual = (int16_t)seg_qual(g); e.mq1 = (uint16_t)mq1;
        e.hash = g.hash; e.rot = g.rotn & 0x3fffffff; e.flags = (g.rotn >> 30) & 1 ? 2u : 0u;
        e.src = seq0 + g.roff;
        s.ins(il - L0, seg_qual(g), mq1, e);
      }
    }
    return 0;
  }
  if (kd == SEG_D) {                                       // (:163-183)
    const int32_t e1 = g.locus + g.len - 1;
    if (g.locus >= J.start && g.locus <= J.stop && e1 >= J.start && e1 <= J.stop) {
      const int32_t dl = g.fin;
      if (dl < J.start) s.edge(r);
      else if (dl >= L0 && dl < L1) {
        IndelRec e;
        e.key = (uint32_t)(((uint64_t)(J.out + (dl - J.start)) << 1) | 1u);
        e.len = g.len; e.qual = (int16_t)seg_qual(g); e.mq1 = (uint16_t)mq1;
        e.hash = g.hash; e.rot = 0; e.flags = (g.rotn >> 30) & 1 ? 2u : 0u;
        e.src = (int64_t)dl - 1;
        s.del(dl - L0, seg_qual(g), mq1, e);
      }
    }
    return 0;
  }
  if (kd == SEG_A) {                                       // slide add (:174-178)
    const int32_t d0 = g.pad, d1 = g.pad + g.len - 1;       // the deletion's gate
    const bool dgate = d0 >= J.start && d0 <= J.stop && d1 >= J.start && d1 <= J.stop;
    const int32_t dl = g.locus, t = g.locus + g.len;
    if (dgate && dl >= J.start && dl <= J.stop && t >= J.start && t <= J.stop && t >= L0 && t < L1) {
      const int32_t q = seg_qual(g);
      if (g.fin >= 0 && q >= cf.min_qual) s.base(t - L0, g.fin, q, mq1);
      s.count_base();
    }
  }
  return 0;
}

// Whether seg_apply would return a valid (base loop) run for g: lets a caller
// reserve run-table space before applying any segment of the read.
PL_HD bool seg_has_run(const KConf& cf, uint32_t info, int32_t L, int32_t L0, int32_t L1, const Seg& g) {
  if (seg_kind(g) != SEG_M || !(info & RI_VALID)) return false;
  int32_t i0 = L0 - g.locus, i1 = g.len;
  if (i0 < 0) i0 = 0;
  if (cf.flank - g.roff > i0) i0 = cf.flank - g.roff;
  if (L1 - g.locus < i1) i1 = L1 - g.locus;
  if (L - cf.flank - g.roff < i1) i1 = L - cf.flank - g.roff;
  return i1 > i0;
}

// ------------------------------------------------------------ column call
struct Col {                 // one PileUp, Long BaseSums widened
  int64_t bc[4], qs[4];
  int32_t mqSum, qSum, deletions, insertions, insQual, delQual;
  PL_HD int64_t count() const { return bc[0] + bc[1] + bc[2] + bc[3]; }
  PL_HD int64_t depth() const { return count() + deletions; }
};

// stable descending order of 0..3 (BaseSum.order, BaseSum.scala:57-60):
// index i's rank is #{j: s[j] > s[i]} + #{j < i: s[j] == s[i]}; written with
// scalar selects so no local array needs addressable memory.
PL_HD void order4(const int64_t s[4], int& o0, int& o1) {
  const int64_t a = s[0], b = s[1], c = s[2], d = s[3];
  const int r0 = (b > a) + (c > a) + (d > a);
  const int r1 = (a > b) + (c > b) + (d > b) + (a == b);
  const int r2 = (a > c) + (b > c) + (d > c) + (a == c) + (b == c);
  o0 = r0 == 0 ? 0 : r1 == 0 ? 1 : r2 == 0 ? 2 : 3;
  o1 = r0 == 1 ? 0 : r1 == 1 ? 1 : r2 == 1 ? 2 : 3;
}

PL_HD int64_t pick4(const int64_t s[4], int i) { return i == 0 ? s[0] : i == 1 ? s[1] : i == 2 ? s[2] : s[3]; }

struct BaseFacts { int base, alt; bool homo; int64_t score, q; };

// the base half of new BaseCall (PileUp.scala:133-150,166-167)
PL_HD BaseFacts base_facts(const Col& c) {
  BaseFacts f;
  const int64_t n = c.count();
  int o0, o1;
  order4(c.qSum > 0 ? c.qs : c.bc, o0, o1);
  f.base = n > 0 ? o0 : 4;
  f.alt = o1;
  const int64_t baseSum = pick4(c.qs, o0), altBaseSum = pick4(c.qs, o1);
  const int64_t total = c.qs[0] + c.qs[1] + c.qs[2] + c.qs[3];
  const int64_t homoScore = baseSum - (total - baseSum);
  const int64_t half = total / 2;
  int64_t d1 = half - baseSum, d2 = half - altBaseSum;
  d1 = d1 < 0 ? -d1 : d1;
  d2 = d2 < 0 ? -d2 : d2;
  const int64_t heteroScore = total - d1 - d2;
  f.homo = homoScore >= heteroScore;
  int64_t dd = homoScore - heteroScore;
  dd = dd < 0 ? -dd : dd;
  // (|homo-hetero| * n) / mqSum and score / n as Long; 32-bit when they fit
  const int64_t num = dd * n;
  if (c.mqSum <= 0) f.score = 0;
  else if ((uint64_t)num <= 0xffffffffull) f.score = (int64_t)((uint32_t)num / (uint32_t)c.mqSum);
  else f.score = num / (int64_t)c.mqSum;
  if (n <= 0) f.q = 0;
  else if ((uint64_t)f.score <= 0xffffffffull && (uint64_t)n <= 0xffffffffull) f.q = (int64_t)((uint32_t)f.score / (uint32_t)n);
  else f.q = f.score / n;
  return f;
}

PL_HD int32_t ins_pct(const Col& c) {                     // PileUp.insPct (:122)
  int32_t a = pct_i(c.insQual, c.mqSum), b = pct_i(c.insertions, (int32_t)c.count());
  return a > b ? a : b;
}
PL_HD int32_t del_pct(const Col& c) {                     // PileUp.delPct (:123)
  int32_t a = pct_i(c.delQual, c.mqSum), b = pct_i(c.deletions, j_add((int32_t)c.count(), c.deletions));
  return a > b ? a : b;
}

// Which indel list BaseCall would consult (insertCall/deletionCall, :183-191)
// and whether hetIndelCall can get past its first test (:213) — if not, the
// answer is ("", true) without looking at the strings.
// returns 0: no indel; 1: insertion list; 2: deletion list
PL_HD int indel_kind_needed(const Col& c, const KConf& cf, int32_t* pct) {
  if (c.insertions > 2 && c.insertions > c.deletions) {
    *pct = ins_pct(c);
    return (c.depth() >= cf.min_min_depth && *pct >= 5) ? 1 : 0;
  }
  if (c.deletions > 2 && c.deletions > c.insertions) {
    *pct = del_pct(c);
    return (c.depth() >= cf.min_min_depth && *pct >= 5) ? 2 : 0;
  }
  return 0;
}

// hetIndelCall after the strings are grouped (PileUp.scala:214-246): n = list
// length, win = count of the majority candidate, L = its length.
// returns 0 none, 1 het, 2 homozygous
PL_HD int het_indel_decide(int32_t n, int32_t win, int32_t L, bool hasN, int32_t pct, const KConf& cf) {
  if (n == 0) return 0;
  if (win < 2 || win <= n / 2) return 0;
  if (hasN) return 0;
  if (cf.old_indel) return (pct >= 33 && pct >= 50 - L) ? 2 : 0;
  int32_t middle = (45 - L) > 10 ? (45 - L) : 10;
  int32_t low = middle / 2;
  int32_t high = middle + middle - low;
  if (pct > high) return 2;
  if (pct >= low) return 1;
  return 0;
}

// pl_call + candidate disposition (GenomeRegion.scala:255-271 assuming
// depth >= minDepth and not deleted; those two tests are applied later).
// indel: 0 none, 1 insertion, 2 deletion; homo_indel as BaseCall.homoIndel.
PL_HD pl_call make_call(const BaseFacts& f, int indel, bool homo_indel, uint8_t refb, const KConf& cf) {
  pl_call out;
  out.base = (uint8_t)f.base;
  out.alt = (uint8_t)f.alt;
  out.q = (int32_t)f.q;
  const bool called = f.base != 4 || indel != 0;
  const bool hc = f.q >= 10;
  uint16_t fl = 0;
  if (f.homo) fl |= PL_CF_HOMO;
  if (hc) fl |= PL_CF_HIGHCONF;
  if (called) fl |= PL_CF_CALLED;
  if (indel) fl |= PL_CF_INDEL;
  if (homo_indel) fl |= PL_CF_HOMOINDEL;
  if (indel == 1) fl |= PL_CF_INSERTION;
  if (indel == 2) fl |= PL_CF_DELETION;
  if (f.score > 0) fl |= PL_CF_SCOREPOS;
  const uint8_t rr = to_upper(refb);
  const uint8_t b = f.base == 4 ? (uint8_t)'N' : acgt_char(f.base);
  int disp = PL_D_NONE;
  if (rr != 'N' && called) {
    if (f.homo && b == rr && hc && !indel) disp = PL_D_CONFIRMED;
    else if (indel == 1 && homo_indel) disp = PL_D_INS;
    else if (indel == 2 && homo_indel) disp = PL_D_DEL;
    else if (b != rr && f.score > 0) {
      if (f.homo) disp = PL_D_SNP;
      else if (cf.fixamb || acgt_char(f.alt) != rr) disp = PL_D_AMB;
    }
  }
  fl |= (uint16_t)(disp << PL_CF_DISP_SHIFT);
  out.flags = fl;
  return out;
}

PL_HD int call_disp(const pl_call& c) { return (c.flags & PL_CF_DISP_MASK) >> PL_CF_DISP_SHIFT; }

// The confirmed count of Pass 1 (GenomeRegion.scala:255-256) is read from a
// per-region histogram of confirmed candidates by depth, whose last bucket
// holds depth 255 and above; a region whose minDepth is above that bucket
// counts its confirmed positions exactly instead, after the deletion pass.
#define PL_CONF_HIST_MAX 255
// Is position g confirmed after Pass 1: a CONFIRMED candidate, not deleted,
// depth >= minDepth, in a region with reads (GenomeRegion.scala:229-256)?
PL_HD bool confirmed_final(const pl_call& c, int64_t depth, int32_t md, int64_t read_count) {
  return read_count != 0 && !(c.flags & PL_CF_DELETED) && depth >= md && call_disp(c) == PL_D_CONFIRMED;
}

// Java Math.round on the double minDepth * meanCoverage (GenomeRegion.scala:221-224)
PL_HD int32_t min_depth_of(const KConf& cf, int64_t base_count, int32_t size) {
  if (cf.min_depth >= 1) return cf.min_depth >= 2147483647.0 ? INT32_MAX : (int32_t)cf.min_depth;   // Double.toInt saturates
  const int64_t cov = round_div_l(base_count, (int64_t)size);
  const double x = cf.min_depth * (double)cov;
  const double f = floor(x);
  int64_t rd = (int64_t)f + ((x - f) >= 0.5 ? 1 : 0);
  int32_t md = (int32_t)rd;
  return md > cf.min_min_depth ? md : cf.min_min_depth;
}

// decode one position's counters from the field-major pileup
PL_HD Col load_col(const int32_t* pile, int64_t G, int64_t g) {
  Col c;
  for (int b = 0; b < 4; b++) {
    c.bc[b] = (int64_t)(uint32_t)pile[(int64_t)(PL_F_COUNT_A + b) * G + g];
    c.qs[b] = (int64_t)pile[(int64_t)(PL_F_QUAL_A + b) * G + g];
  }
  c.mqSum = pile[(int64_t)PL_F_MQSUM * G + g];
  c.qSum = pile[(int64_t)PL_F_QSUM * G + g];
  c.deletions = pile[(int64_t)PL_F_DELETIONS * G + g];
  c.insertions = pile[(int64_t)PL_F_INSERTIONS * G + g];
  c.insQual = pile[(int64_t)PL_F_INSQUAL * G + g];
  c.delQual = pile[(int64_t)PL_F_DELQUAL * G + g];
  return c;
}

// ------------------------------------------------------------ indel majority
// Two records name the same string?  Deletions at one position differ only
// by length (the string is the reference at that position); insertions are
// compared by length, hash, then byte by byte.
PL_HD uint8_t rec_char(const RSDev& R, const IndelRec& e, int32_t j) {
  int32_t oi = (j - e.rot) % e.len;
  if (oi < 0) oi += e.len;
  int64_t q = e.src + oi;
  if (e.flags & 1u) { uint8_t b = R.nib[q >> 1]; return nib_char((q & 1) ? (b & 15) : (b >> 4)); }
  return acgt_char((R.bases2[q >> 2] >> (2 * (q & 3))) & 3);
}
PL_HD bool rec_same(const RSDev& R, const IndelRec& x, const IndelRec& y) {
  if (x.len != y.len) return false;
  if ((x.key & 1u) == 1u) return true;
  if (x.hash != y.hash) return false;
  for (int32_t j = 0; j < x.len; j++) if (rec_char(R, x, j) != rec_char(R, y, j)) return false;
  return true;
}

// Exact strict-majority test over one (position, kind) group of records:
// a string with count > n/2 exists iff the Boyer-Moore candidate has it, so
// the outcome is that of map.toSeq.sortBy(_._2).last with the strict-majority
// test of hetIndelCall, independent of order.
struct Majority { int32_t n, win; int64_t rec; };
PL_HD Majority group_majority(const RSDev& R, const IndelRec* recs, const uint32_t* idx, int64_t a, int64_t b) {
  Majority m;
  m.n = (int32_t)(b - a);
  m.win = 0;
  m.rec = -1;
  if (b <= a) return m;
  int64_t cand = idx[a];
  int32_t cnt = 0;
  for (int64_t i = a; i < b; i++) {
    const IndelRec& e = recs[idx[i]];
    if (cnt == 0) { cand = idx[i]; cnt = 1; }
    else if (rec_same(R, e, recs[cand])) cnt++;
    else cnt--;
  }
  int32_t c = 0;
  for (int64_t i = a; i < b; i++) if (rec_same(R, recs[idx[i]], recs[cand])) c++;
  m.win = c;
  m.rec = cand;
  return m;
}

PL_HD int64_t lower_bound_u32(const uint32_t* k, int64_t n, uint32_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) { int64_t m = (lo + hi) >> 1; if (k[m] < x) lo = m + 1; else hi = m; }
  return lo;
}

// ------------------------------------------------------------ window flush
// Window accumulators use packed 64-bit adds.  A per-base word holds
// count * 2^40 + sum(qual * mq1) (mod 2^64): the sum is BaseSum's Long
// qualSum (BaseSum.scala:24), kept exact inside +-2^39, and the count stays
// below 2^24; one 64-bit atomic updates both.  The other words hold a pair
// (lo, hi) of Java Ints as lo + hi * 2^32: lo (mqSum, insQual, delQual: at
// most 256 per base) must stay inside int32, hi then wraps like a Java Int.
PL_HD void unpack_pair(uint64_t x, int32_t* lo, int32_t* hi) {
  int32_t l = (int32_t)(uint32_t)x;
  *lo = l;
  *hi = (int32_t)(uint32_t)((x - (uint64_t)(int64_t)l) >> 32);
}
#define PL_BASE_SHIFT 40
PL_HD void unpack_base(uint64_t x, int64_t* qual_sum, uint32_t* count) {
  const int64_t l = (int64_t)(x << (64 - PL_BASE_SHIFT)) >> (64 - PL_BASE_SHIFT);   // low 40 bits, signed
  *qual_sum = l;
  *count = (uint32_t)((x - (uint64_t)l) >> PL_BASE_SHIFT);
}
PL_HD uint64_t pack_base(int32_t qual, int32_t mq1) {        // count + qual*mq1
  return (1ull << PL_BASE_SHIFT) + (uint64_t)(int64_t)j_mul(qual, mq1);
}
PL_HD uint64_t pack_mq_q(int32_t mq1, int32_t qual) {        // mqSum += mq1, qSum += qual
  return (uint64_t)(int64_t)mq1 + ((uint64_t)(int64_t)qual << 32);
}

// Bases (plus deletions) one column may hold: beyond it the packed mqSum
// (256 per base at MAPQ 255) could carry into qSum, and Pilon's own Int
// mqSum wraps; refused with PL_E_OVERFLOW.
#define PL_DEPTH_LIMIT (1 << 23)

// A column whose Long qualSum of some base leaves int32 ("wide"): the
// pileup planes hold the low 32 bits (Long.toInt), PL_CF_WIDE marks its
// call, and the four exact sums are kept beside the batch in one list per
// window.  Every pass that reads qualSum again patches them in.
struct WideCol {
  uint32_t gpos;
  int32_t pad;
  int64_t qs[4];
};

// The exact qualSum of a wide column (PL_CF_WIDE) into c: the entry of
// batch position g in its window's list (win_wide: first entry and count
// per window; windows of W positions from each region's first).
PL_HD void wide_patch(Col& c, const WideCol* wide, uint32_t wide_cap, const int32_t* win_wide, const JobDev& J,
                      int32_t W, int64_t g) {
  const int32_t w = J.win0 + (int32_t)((g - J.out) / W);
  const int64_t off = win_wide[2 * w];
  int64_t n = win_wide[2 * w + 1];
  if (off + n > (int64_t)wide_cap) n = (int64_t)wide_cap - off;   // an overfull list: the batch runs again
  for (int64_t k = 0; k < n; k++)
    if (wide[off + k].gpos == (uint32_t)g) {
      for (int b = 0; b < 4; b++) c.qs[b] = wide[off + k].qs[b];
      return;
    }
}

struct PosOut {
  int32_t f[PL_NFIELDS];
  pl_call call;
  bool defer;        // needs the indel strings (indel pass)
  bool overflow;
  bool wide;         // some qualSum leaves int32 (exact sums in qs)
  int64_t qs[4];
};

// One column at the end of a window: counters, physCov (PileUpRegion.scala:
// 90-100 from the inclusive sums), and BaseCall unless the indel strings are
// needed.
PL_HD void finalize_position(const uint64_t bs[4], uint64_t mqq, uint64_t ins, uint64_t del,
                             uint32_t bad, uint32_t clips, int32_t pc, int32_t is, uint8_t refb,
                             const KConf& cf, PosOut& o) {
  Col c;
  o.overflow = false;
  o.wide = false;
  uint32_t depth = 0;
  for (int b = 0; b < 4; b++) {
    int64_t qs;
    uint32_t n;
    unpack_base(bs[b], &qs, &n);
    c.bc[b] = (int64_t)n;
    c.qs[b] = qs;
    o.qs[b] = qs;
    o.f[PL_F_COUNT_A + b] = (int32_t)n;
    o.f[PL_F_QUAL_A + b] = (int32_t)qs;   // Long.toInt; exact value in the wide list
    if (qs != (int64_t)(int32_t)qs) o.wide = true;
    depth += n;
  }
  unpack_pair(mqq, &c.mqSum, &c.qSum);
  unpack_pair(ins, &c.insQual, &c.insertions);
  unpack_pair(del, &c.delQual, &c.deletions);
  if (depth + (uint32_t)c.deletions >= PL_DEPTH_LIMIT || (uint32_t)c.insertions >= PL_DEPTH_LIMIT) o.overflow = true;
  o.f[PL_F_MQSUM] = c.mqSum;
  o.f[PL_F_QSUM] = c.qSum;
  o.f[PL_F_PHYSCOV] = pc;
  o.f[PL_F_INSERTSIZE] = pc > 0 ? j_div(is, pc) : is;
  o.f[PL_F_BADPAIR] = (int32_t)bad;
  o.f[PL_F_DELETIONS] = c.deletions;
  o.f[PL_F_DELQUAL] = c.delQual;
  o.f[PL_F_INSERTIONS] = c.insertions;
  o.f[PL_F_INSQUAL] = c.insQual;
  o.f[PL_F_CLIPS] = (int32_t)clips;
  BaseFacts bf = base_facts(c);
  int32_t pct;
  int need = indel_kind_needed(c, cf, &pct);
  o.defer = need != 0;
  o.call = make_call(bf, 0, true, refb, cf);
}

// The full BaseCall of one position from the pileup and its sorted evidence
// records (indel pass and recomputation after Pass-1 mutations).
PL_HD pl_call full_call(const RSDev* rsets, const JobDev& J, const Col& c, uint8_t refb,
                        const IndelRec* recs, const uint32_t* rkey, const uint32_t* rval, int64_t nrec,
                        uint32_t gpos, const KConf& cf, int32_t* indel_len, int64_t* win_rec) {
  BaseFacts bf = base_facts(c);
  int32_t pct = 0;
  int need = indel_kind_needed(c, cf, &pct);
  int indel = 0;
  bool homo = true;
  *indel_len = 0;
  *win_rec = -1;
  if (need) {
    const uint32_t key = (gpos << 1) | (uint32_t)(need - 1);
    const int64_t a = lower_bound_u32(rkey, nrec, key);
    const int64_t b = lower_bound_u32(rkey, nrec, key + 1);
    Majority m = group_majority(rsets[J.rs], recs, rval, a, b);
    if (m.rec >= 0) {
      const IndelRec& e = recs[m.rec];
      int d = het_indel_decide(m.n, m.win, e.len, (e.flags & 2u) != 0, pct, cf);
      if (d) {
        indel = need;
        homo = d == 2;
        *indel_len = e.len;
        *win_rec = m.rec;
      }
    }
  }
  return make_call(bf, indel, homo, refb, cf);
}

// GenomeRegion.postProcess Pass 1 for one region (GenomeRegion.scala:221-272),
// run after every column has its candidate disposition: minDepth, the
// confirmed count, and the left-to-right deletion pass that marks
// deleted(i+j) and adds pu.deletions into those pileups (:259-264).
// cand[c_lo..c_hi) are the region's SNP/AMB/INS/DEL candidates sorted by
// position; kind is set to 0 for those that do not become changes.
PL_HD void finalize_job(const KConf& cf, const JobDev& J, JobStat& st, int32_t* pile, int64_t G,
                        pl_call* calls, Cand* cand, int64_t c_lo, int64_t c_hi,
                        uint32_t* mutated, uint32_t* mut_n, uint32_t mut_cap, int32_t* overflow) {
  const int32_t md = min_depth_of(cf, st.base_count, J.size);
  st.min_depth = md;
  // confirmed positions by depth, from the histogram when minDepth falls
  // inside it; otherwise counted exactly after the pass (confirmed_exact)
  int64_t nconf = 0;
  if (md <= PL_CONF_HIST_MAX) for (int d = md; d < 256; d++) nconf += st.hist[d];
  int64_t delEnd = -1;
  for (int64_t k = c_lo; k < c_hi; k++) {
    Cand& c = cand[k];
    const int64_t g = c.gpos;
    if (st.read_count == 0 || g <= delEnd || c.depth < md) { c.kind = 0; continue; }
    if (c.kind == PL_D_DEL) {
      const int32_t dcount = pile[(int64_t)PL_F_DELETIONS * G + g];
      for (int32_t j = 1; j < c.len; j++) {
        const int64_t gj = g + j;
        int64_t dep = 0;
        for (int b = 0; b < 4; b++) dep += (int64_t)(uint32_t)pile[(int64_t)(PL_F_COUNT_A + b) * G + gj];
        dep += pile[(int64_t)PL_F_DELETIONS * G + gj];
        if (md <= PL_CONF_HIST_MAX && call_disp(calls[gj]) == PL_D_CONFIRMED && dep >= md && nconf > 0) nconf--;
        pile[(int64_t)PL_F_DELETIONS * G + gj] = j_add(pile[(int64_t)PL_F_DELETIONS * G + gj], dcount);
        calls[gj].flags |= PL_CF_DELETED;
        uint32_t slot = pl_bump(mut_n);
        if (slot < mut_cap) mutated[slot] = (uint32_t)gj;
        else *overflow = 1;
      }
      delEnd = g + c.len - 1;
    }
  }
  st.n_confirmed = (int32_t)nconf;
}
