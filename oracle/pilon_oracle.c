/*
 * pilon_oracle.c — CPU restatement of Pilon 1.24's pileup-construction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU engine is
 * compared against; it is never linked into libpilonhip.so and nothing in the
 * product calls it.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py load it.
 *
 * It follows the Scala sources line by line, one read and one position at a
 * time, with Java's integer semantics (Int wraps at 32 bits, Long at 64,
 * division truncates).  Every function names the reference lines it restates.
 * Reference: broadinstitute/pilon v1.24, src/main/scala/org/broadinstitute/pilon/
 *
 * Pins (tests/test_oracle_kat.py): hand-derived known answers per addRead
 * rule, the physCov closed form, the flank-0 count identity, Math.round
 * corner cases, stable order ties, BaseCall scores, the hetIndelCall
 * threshold table, the oldIndel branch, strict majority and N rejection, the
 * AMB disposition and its change line with and without --iupac, planted
 * variants fixed.
 * Further pins: the VCF records (tests/test_oracle_vcf.py: hand-derived
 * lines, "%.2f" of the Float AF against decimal rounding half up), the
 * tracks (tests/test_oracle_tracks.py: smooth's closed forms and edges,
 * computeGc against a brute force, delta/dip/copy number from fragment
 * coverage rebuilt from the final pileup), Pilon.bed
 * (tests/test_oracle_bed.py: summaryRegions against run grouping, every
 * region's lines against a second derivation), the assembler's K-mer graph
 * (tests/test_oracle_kmer.py: against a dict-of-pileups restatement).
 * Not restated: Assembler.addRead's prune at its 100,000th read (the K-mer
 * graph refuses such inputs, PL_E_INVAL, as the device pass does).
 * Unreachable here: the overlap branch of fixFixList (GenomeRegion.scala:
 * 557-593).  Pass 1 adds at most one change per locus and marks the loci a
 * deletion covers as deleted, so no two fixes it makes overlap.
 * Parity unpinned (no independent check of this reading of the reference):
 * the htsjdk decoding of non-ACGT SEQ codes (the host packer's side of the
 * boundary).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <math.h>
#include "../include/pilon_hip.h"

/* ---------------------------------------------------------- Java ints */
static inline int32_t jadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t jsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t jmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
static inline int32_t jdiv(int32_t a, int32_t b) {            /* b != 0 */
  if (a == INT32_MIN && b == -1) return INT32_MIN;
  return a / b;
}
static inline int64_t ldiv64(int64_t a, int64_t b) {
  if (a == INT64_MIN && b == -1) return INT64_MIN;
  return a / b;
}

/* Utils.roundDiv / Utils.pct, Int and Long overloads (Utils.scala:23-26). */
static int32_t roundDivI(int32_t n, int32_t d) { return d > 0 ? jdiv(jadd(n, d / 2), d) : 0; }
static int64_t roundDivL(int64_t n, int64_t d) { return d > 0 ? ldiv64(n + d / 2, d) : 0; }
static int32_t pctI(int32_t n, int32_t d) { return roundDivI(jmul(100, n), d); }

/* java.lang.Math.round(double): the nearest long, ties toward +infinity
 * (its javadoc).  floor(x) + [x - floor(x) >= 1/2]; the subtraction is exact
 * for every finite double. */
int64_t orc_java_round(double x) {
  if (x != x) return 0;
  if (x >= 9.2233720368547758e18) return INT64_MAX;
  if (x <= -9.2233720368547758e18) return INT64_MIN;
  double f = floor(x);
  return (int64_t)f + ((x - f) >= 0.5 ? 1 : 0);
}

/* ------------------------------------------------------- strings/lists */
typedef struct { const uint8_t* s; int32_t len; } ostr;
typedef struct { ostr* v; int32_t n, cap; } olist;

static void olist_push(olist* l, const uint8_t* s, int32_t len) {
  if (l->n == l->cap) {
    l->cap = l->cap ? 2 * l->cap : 4;
    l->v = (ostr*)realloc(l->v, sizeof(ostr) * (size_t)l->cap);
  }
  l->v[l->n].s = s;
  l->v[l->n].len = len;
  l->n++;
}

/* byte arena for indel strings (they live until the region is freed) */
typedef struct arena_blk { struct arena_blk* next; size_t used, cap; uint8_t data[]; } arena_blk;
typedef struct { arena_blk* head; } arena;
static uint8_t* arena_alloc(arena* a, size_t n) {
  if (!a->head || a->head->used + n > a->head->cap) {
    size_t cap = n > (1u << 20) ? n : (1u << 20);
    arena_blk* b = (arena_blk*)malloc(sizeof(arena_blk) + cap);
    b->next = a->head; b->used = 0; b->cap = cap; a->head = b;
  }
  uint8_t* p = a->head->data + a->head->used;
  a->head->used += n;
  return p;
}
static void arena_free(arena* a) {
  while (a->head) { arena_blk* n = a->head->next; free(a->head); a->head = n; }
}

/* ------------------------------------------------------------- PileUp */
/* PileUp fields (PileUp.scala:26-41); BaseSum sums are Long (BaseSum.scala:24). */
typedef struct {
  int64_t bc[4], qs[4];
  int32_t mqSum, qSum, physCov, insertSize, badPair, deletions, delQual,
          insertions, insQual, clips;
  olist ins, del;
} PileUp;

static int baseIndex(uint8_t c) {                     /* PileUp.scala:46-52 */
  switch (c) { case 'A': return 0; case 'C': return 1; case 'G': return 2; case 'T': return 3; }
  return -1;
}

static int64_t pu_count(const PileUp* p) { return p->bc[0] + p->bc[1] + p->bc[2] + p->bc[3]; }   /* :43 */
static int64_t pu_depth(const PileUp* p) { return pu_count(p) + p->deletions; }                  /* :44 */

/* PileUp.add (PileUp.scala:75-84) */
static void pu_add(PileUp* p, uint8_t base, int32_t qual, int32_t mq, const pl_config* cfg) {
  int bi = baseIndex(base);
  if (bi >= 0 && qual >= cfg->min_qual) {
    int32_t mq1 = jadd(mq, 1);
    p->bc[bi] += 1;
    p->qs[bi] += (int64_t)jmul(qual, mq1);
    p->mqSum = jadd(p->mqSum, mq1);
    p->qSum = jadd(p->qSum, qual);
  }
}
/* PileUp.addInsertion (PileUp.scala:98-105): mqSum untouched. */
static void pu_addInsertion(PileUp* p, const uint8_t* s, int32_t len, int32_t qual, int32_t mq) {
  int32_t mq1 = jadd(mq, 1);
  p->insQual = jadd(p->insQual, mq1);
  p->qSum = jadd(p->qSum, qual);
  olist_push(&p->ins, s, len);
  p->insertions = jadd(p->insertions, 1);
}
/* PileUp.addDeletion (PileUp.scala:107-114) */
static void pu_addDeletion(PileUp* p, const uint8_t* s, int32_t len, int32_t qual, int32_t mq) {
  int32_t mq1 = jadd(mq, 1);
  p->mqSum = jadd(p->mqSum, mq1);
  p->delQual = jadd(p->delQual, mq1);
  p->qSum = jadd(p->qSum, qual);
  olist_push(&p->del, s, len);
  p->deletions = jadd(p->deletions, 1);
}

/* BaseSum.order: stable descending sort of 0..3 (BaseSum.scala:57-60). */
static void basesum_order(const int64_t s[4], int ord[4]) {
  for (int i = 0; i < 4; i++) ord[i] = i;
  for (int i = 1; i < 4; i++) {            /* insertion sort, strict > keeps ties in index order */
    int x = ord[i], j = i - 1;
    while (j >= 0 && s[x] > s[ord[j]]) { ord[j + 1] = ord[j]; j--; }
    ord[j + 1] = x;
  }
}

/* --------------------------------------------------------- BaseCall */
typedef struct {
  int base;              /* 0..3, 4 = 'N' */
  int bI;                /* order(0): the index of baseSum even when base is 'N' */
  int alt;
  int homo;
  int64_t score, q;
  int highConf;
  const uint8_t* indelStr; int32_t indelLen;
  int isIns, isDel, indel, homoIndel, called;
} BaseCall;

static int cmp_ostr(const void* a, const void* b) {
  const ostr* x = (const ostr*)a; const ostr* y = (const ostr*)b;
  if (x->len != y->len) return x->len < y->len ? -1 : 1;
  return memcmp(x->s, y->s, (size_t)x->len);
}

/* BaseCall.hetIndelCall (PileUp.scala:209-247) incl. the oldIndel branch. */
static void hetIndelCall(const PileUp* p, const olist* l, int32_t pct, const pl_config* cfg,
                         const uint8_t** win, int32_t* winLen, int* homo) {
  *win = NULL; *winLen = 0; *homo = 1;
  if (pu_depth(p) < cfg->min_min_depth || pct < 5 || l->n == 0) return;
  /* map(indelStr) += 1 ; winner = map.toSeq.sortBy(_._2).last */
  ostr* v = (ostr*)malloc(sizeof(ostr) * (size_t)l->n);
  memcpy(v, l->v, sizeof(ostr) * (size_t)l->n);
  qsort(v, (size_t)l->n, sizeof(ostr), cmp_ostr);
  int32_t best = 0; ostr bestS = {0, 0};
  for (int32_t i = 0; i < l->n;) {
    int32_t j = i + 1;
    while (j < l->n && cmp_ostr(&v[i], &v[j]) == 0) j++;
    if (j - i >= best) { best = j - i; bestS = v[i]; }
    i = j;
  }
  free(v);
  if (best < 2 || best <= l->n / 2) return;
  for (int32_t i = 0; i < bestS.len; i++) if (bestS.s[i] == 'N') return;
  int32_t L = bestS.len;
  if (cfg->old_indel) {
    if (pct >= 33 && pct >= 50 - L) { *win = bestS.s; *winLen = L; *homo = 1; }
    return;
  }
  int32_t middle = (45 - L) > 10 ? (45 - L) : 10;
  int32_t low = middle / 2;
  int32_t high = middle + middle - low;
  if (pct > high) { *win = bestS.s; *winLen = L; *homo = 1; }
  else if (pct >= low) { *win = bestS.s; *winLen = L; *homo = 0; }
}

/* PileUp.insPct / delPct (PileUp.scala:122-123) */
static int32_t insPct(const PileUp* p) {
  int32_t a = pctI(p->insQual, p->mqSum), b = pctI(p->insertions, (int32_t)pu_count(p));
  return a > b ? a : b;
}
static int32_t delPct(const PileUp* p) {
  int32_t a = pctI(p->delQual, p->mqSum);
  int32_t b = pctI(p->deletions, jadd((int32_t)pu_count(p), p->deletions));
  return a > b ? a : b;
}

/* new BaseCall (PileUp.scala:132-191) */
static void baseCall(const PileUp* p, const pl_config* cfg, BaseCall* bc) {
  memset(bc, 0, sizeof(*bc));
  int64_t n = pu_count(p);
  int ord[4];
  basesum_order(p->qSum > 0 ? p->qs : p->bc, ord);
  bc->base = n > 0 ? ord[0] : 4;
  int bI = ord[0], aI = ord[1];
  bc->bI = bI;
  bc->alt = aI;
  int64_t baseSum = p->qs[bI], altBaseSum = p->qs[aI];
  int64_t total = p->qs[0] + p->qs[1] + p->qs[2] + p->qs[3];
  int64_t homoScore = baseSum - (total - baseSum);
  int64_t halfTotal = ldiv64(total, 2);
  int64_t d1 = halfTotal - baseSum, d2 = halfTotal - altBaseSum;
  int64_t heteroScore = total - (d1 < 0 ? -d1 : d1) - (d2 < 0 ? -d2 : d2);
  bc->homo = homoScore >= heteroScore;
  int64_t dd = homoScore - heteroScore; if (dd < 0) dd = -dd;
  bc->score = p->mqSum > 0 ? ldiv64(dd * n, (int64_t)p->mqSum) : 0;
  bc->q = n > 0 ? ldiv64(bc->score, n) : 0;
  bc->highConf = bc->q >= 10;
  /* insertCall / deletionCall (PileUp.scala:151-162,183-191) */
  const uint8_t* s = NULL; int32_t L = 0; int h = 1;
  if (p->insertions > 2 && p->insertions > p->deletions)
    hetIndelCall(p, &p->ins, insPct(p), cfg, &s, &L, &h);
  if (s) { bc->isIns = 1; bc->indel = 1; bc->homoIndel = h; bc->indelStr = s; bc->indelLen = L; }
  else {
    h = 1;
    if (p->deletions > 2 && p->deletions > p->insertions)
      hetIndelCall(p, &p->del, delPct(p), cfg, &s, &L, &h);
    if (s) { bc->isDel = 1; bc->indel = 1; bc->homoIndel = h; bc->indelStr = s; bc->indelLen = L; }
    else { bc->homoIndel = 1; }
  }
  bc->called = (bc->base != 4) || bc->indel;
}

/* ------------------------------------------------------- PileUpRegion */
typedef struct {
  int32_t start, stop, size;
  PileUp* pu;
  int64_t baseCount;
  int32_t readCount;
  int32_t physCovStart, insertSizeStart;
  int64_t unknownOps;
  int64_t alignedBases;
  arena ar;
  int edge;  int64_t edgeRead; const char* edgeWhy;
} PRegion;

static inline int inRegion(const PRegion* R, int64_t l) { return l >= R->start && l <= R->stop; }   /* Region.scala:23 */
static inline int beforeRegion(const PRegion* R, int64_t l) { return l < R->start; }             /* :24 */

/* PileUpRegion.add / remove (PileUpRegion.scala:38-58) */
static void pr_add(PRegion* R, int64_t locus, uint8_t base, int32_t qual, int32_t mq, int pair,
                   const pl_config* cfg) {
  if (inRegion(R, locus)) {
    PileUp* p = &R->pu[locus - R->start];
    if (pair) { pu_add(p, base, qual, mq, cfg); R->baseCount += 1; }
    else p->badPair = jadd(p->badPair, 1);
  }
}
static void pr_remove(PRegion* R, int64_t locus, int pair) {
  if (inRegion(R, locus)) {
    if (!pair) R->pu[locus - R->start].badPair = jsub(R->pu[locus - R->start].badPair, 1);
  }
}

/* PileUpRegion.physCovIncr (PileUpRegion.scala:62-88) */
static int32_t pr_physCovIncr(PRegion* R, int32_t aStart, int32_t aEnd, int32_t iSize, int paired, int valid) {
  if (!valid || (paired && iSize <= 0)) return 0;
  int32_t start, end;
  if (!paired) { start = aStart < aEnd ? aStart : aEnd; end = aStart > aEnd ? aStart : aEnd; }
  else if (iSize > 0) { start = aStart; end = jadd(aStart, iSize); }
  else { start = jadd(jadd(aEnd, 1), iSize); end = jadd(aEnd, 1); }
  int32_t insertSize = jsub(end, start);
  if (inRegion(R, start)) {
    PileUp* p = &R->pu[start - R->start];
    p->physCov = jadd(p->physCov, 1);
    p->insertSize = jadd(p->insertSize, insertSize);
  } else if (beforeRegion(R, start) && !beforeRegion(R, end)) {
    R->physCovStart = jadd(R->physCovStart, 1);
    R->insertSizeStart = jadd(R->insertSizeStart, insertSize);
  }
  if (inRegion(R, end)) {
    PileUp* p = &R->pu[end - R->start];
    p->physCov = jsub(p->physCov, 1);
    p->insertSize = jsub(p->insertSize, insertSize);
  }
  return insertSize;
}

/* PileUpRegion.computePhysCov (PileUpRegion.scala:90-100) */
static void pr_computePhysCov(PRegion* R) {
  R->pu[0].physCov = jadd(R->pu[0].physCov, R->physCovStart);
  R->pu[0].insertSize = jadd(R->pu[0].insertSize, R->insertSizeStart);
  for (int32_t i = 1; i < R->size; i++) {
    R->pu[i].physCov = jadd(R->pu[i].physCov, R->pu[i - 1].physCov);
    R->pu[i].insertSize = jadd(R->pu[i].insertSize, R->pu[i - 1].insertSize);
  }
  for (int32_t i = 0; i < R->size; i++)
    if (R->pu[i].physCov > 0) R->pu[i].insertSize = jdiv(R->pu[i].insertSize, R->pu[i].physCov);
}

/* A decoded SAMRecord view. */
typedef struct {
  int32_t length;           /* getReadLength */
  const uint8_t* bases;     /* getReadBases (ASCII) */
  const int8_t* quals;      /* getBaseQualities or the defaultQual fill */
  int32_t mq, flag, tlen, aStart, aEnd;
  const uint32_t* cigar; int64_t ncigar;
  int64_t index;
} ReadView;

static const char NIB[] = "=ACMGRSVTWYHKDBN";
static int consumesRead(int op) { return op == 0 || op == 1 || op == 4 || op == 7 || op == 8; }
static int consumesRef(int op) { return op == 0 || op == 2 || op == 3 || op == 7 || op == 8; }

#define EDGE(R, r, why) do { if (!(R)->edge) { (R)->edge = 1; (R)->edgeRead = (r)->index; (R)->edgeWhy = (why); } return; } while (0)

/* addRead's homoRun (PileUpRegion.scala:120-126): the run of equal bytes of
 * the contig from 0-based index i0 to the contig's end.  refBases(i0) out of
 * range throws (*oob). */
static int32_t pr_homoRun(const uint8_t* refBases, int64_t contig_len, int64_t i0, int* oob) {
  if (i0 < 0 || i0 >= contig_len) { *oob = 1; return 0; }
  for (int64_t i = i0 + 1; i < contig_len; i++)
    if (refBases[i] != refBases[i0]) return (int32_t)(i - i0);
  return (int32_t)(contig_len - i0);
}

/* addRead's nanoporeExclude (PileUpRegion.scala:128-134): i0 is a region
 * index (index(locus)) but indexes the whole contig's bytes; the region test
 * is on locus(i0 -+ 2). */
static int pr_nanoporeExclude(const PRegion* R, const uint8_t* refBases, int64_t i0) {
  return inRegion(R, (int64_t)R->start + i0 - 2) && inRegion(R, (int64_t)R->start + i0 + 2) &&
         refBases[i0 - 2] == 'C' && refBases[i0 - 1] == 'C' && refBases[i0 + 1] == 'G' && refBases[i0 + 2] == 'G';
}

#define LR_NANOPORE 1   /* BamFile.nanoporeLongRead (BamFile.scala:34) */
#define LR_PACBIO 2     /* BamFile.pacbioLongRead (:35) */

/* PileUpRegion.addRead (PileUpRegion.scala:102-220); longRead is the BAM's
 * longReadType (BamFile.scala:43-47): 0, LR_NANOPORE or LR_PACBIO. */
static void pr_addRead(PRegion* R, const ReadView* r, const uint8_t* refBases, int64_t contig_len, int32_t longRead,
                       const pl_config* cfg) {
  const int32_t length = r->length;
  const uint8_t* bases = r->bases;
  const int32_t mq = r->mq;
  const int paired = (r->flag & 0x1) != 0;
  const int valid = (mq >= cfg->min_mq) &&
      (!paired || (((r->flag & 0x2) != 0) && ((r->flag & PL_RF_MATE_SAME_REF) != 0)));
  const int32_t insert = r->tlen, aStart = r->aStart, aEnd = r->aEnd;
  const int8_t* quals = r->quals;
  const int32_t flank = cfg->flank;
#define TRUSTED(o) ((o) >= flank && jsub(length, flank) > (o))

  int32_t clipped = 0;
  for (int64_t k = 0; k < r->ncigar; k++)
    if ((r->cigar[k] & 0xf) == 4) clipped = jadd(clipped, (int32_t)(r->cigar[k] >> 4));
  int32_t adjMq = roundDivI(jmul(mq, jsub(length, clipped)), length);      /* :139-141 */
  int32_t indelMq = longRead > 0 ? (adjMq < 8 ? adjMq : 8) : adjMq;       /* :142 */
  int oob = 0;

  int32_t readOffset = 0, refOffset = 0;
  for (int64_t k = 0; k < r->ncigar; k++) {
    int32_t len = (int32_t)(r->cigar[k] >> 4);
    int op = (int)(r->cigar[k] & 0xf);
    int32_t locus = jadd(aStart, refOffset);
    if (op == 1) {                                                          /* I :150-162 */
      int32_t iloc = locus;
      if (valid && TRUSTED(readOffset) && inRegion(R, iloc)) {
        uint8_t* ins = arena_alloc(&R->ar, (size_t)len + 1);
        memcpy(ins, bases + readOffset, (size_t)len);
        while (iloc > 1 && refBases[iloc - 2] == ins[len - 1]) {
          iloc -= 1;
          uint8_t last = ins[len - 1];                 /* slice(len-1,len) ++ slice(0,len-1) */
          memmove(ins + 1, ins, (size_t)len - 1);
          ins[0] = last;
        }
        /* :160: a long read's insertion at the start of a homopolymer of 4 is dropped */
        if (longRead > 0) {
          int32_t run = pr_homoRun(refBases, contig_len, iloc, &oob);
          if (oob) EDGE(R, r, "homoRun past the contig end");
          if (run >= 4) goto next_op;
        }
        if (iloc < R->start) EDGE(R, r, "insertion normalised before region start");
        pu_addInsertion(&R->pu[iloc - R->start], ins, len, quals[readOffset], indelMq);
      }
    } else if (op == 2) {                                                   /* D :163-183 */
      int32_t dloc = locus, rloc = readOffset;
      if (valid && TRUSTED(readOffset) && inRegion(R, dloc) && inRegion(R, (int64_t)dloc + len - 1)) {
        while (dloc > 1 && rloc > 0 && refBases[dloc - 2] == refBases[dloc + len - 2]) {
          dloc -= 1;
          rloc -= 1;
          uint8_t base = bases[rloc];
          int32_t qual = quals[rloc];
          if (TRUSTED(rloc) && inRegion(R, dloc)) {
            pr_remove(R, dloc, valid);
            if (inRegion(R, (int64_t)dloc + len)) pr_add(R, (int64_t)dloc + len, base, qual, adjMq, valid, cfg);
          }
        }
        /* :180-182: homoRun and nanoporeExclude of index(dloc), a region
           index read as a contig index; a negative one throws */
        if (longRead > 0) {
          int32_t run = pr_homoRun(refBases, contig_len, (int64_t)dloc - R->start, &oob);
          if (oob) EDGE(R, r, "deletion normalised before region start");
          if (run >= 4) goto next_op;
          if (longRead == LR_NANOPORE && pr_nanoporeExclude(R, refBases, (int64_t)dloc - R->start)) goto next_op;
        }
        if (dloc < R->start) EDGE(R, r, "deletion normalised before region start");
        pu_addDeletion(&R->pu[dloc - R->start], refBases + (dloc - 1), len, quals[readOffset], indelMq);
      }
    } else if (op == 0 || op == 7 || op == 8) {                             /* M/=/X :184-193 */
      for (int32_t i = 0; i < len; i++) {
        int32_t rOff = readOffset + i;
        if (TRUSTED(rOff)) {
          int64_t locusPlus = (int64_t)locus + i;
          /* :190: a nanopore base inside CC_GG piles with quality 0 */
          int32_t qual = (longRead == LR_NANOPORE && pr_nanoporeExclude(R, refBases, locusPlus - R->start)) ? 0 : quals[rOff];
          pr_add(R, locusPlus, bases[rOff], qual, adjMq, valid, cfg);
        }
      }
    } else if (op == 4) {                                                   /* S :194-206 */
      int32_t clipStart = readOffset == 0 ? jsub(locus, len) : locus;
      int32_t clipEnd = jsub(jadd(clipStart, len), 1);
      if (inRegion(R, clipStart)) R->pu[clipStart - R->start].clips = jadd(R->pu[clipStart - R->start].clips, 1);
      if (inRegion(R, clipEnd)) R->pu[clipEnd - R->start].clips = jadd(R->pu[clipEnd - R->start].clips, 1);
      for (int32_t i = 0; i < len; i++) {
        int32_t rOff = readOffset + i;
        if (rOff >= length) EDGE(R, r, "soft clip beyond read bases");
        int64_t locusPlus = (int64_t)clipStart + i;
        if (inRegion(R, locusPlus)) pr_add(R, locusPlus, bases[rOff], quals[rOff], adjMq, 0, cfg);
      }
    } else if (op == 5 || op == 3) {                                        /* H, N :207-210 */
    } else {
      R->unknownOps++;                                                      /* :211-212 */
    }
  next_op:
    if (consumesRead(op)) readOffset = jadd(readOffset, len);
    if (consumesRef(op)) refOffset = jadd(refOffset, len);
  }
  R->readCount = jadd(R->readCount, 1);
  pr_physCovIncr(R, aStart, aEnd, insert, paired, valid);
#undef TRUSTED
}

/* ------------------------------------------------------- read access */
typedef struct {
  const pl_read_soa* s;
  uint8_t* bases; int8_t* quals; int32_t cap;
} Decoder;

/* htsjdk getAlignmentEnd: 0 for an unmapped read, else start + ref length - 1. */
static int32_t alignment_end(const pl_read_soa* s, int64_t i) {
  if (s->flag[i] & 0x4) return 0;
  int32_t rl = 0;
  for (int64_t k = s->cigar_off[i]; k < s->cigar_off[i + 1]; k++)
    if (consumesRef((int)(s->cigar[k] & 0xf))) rl += (int32_t)(s->cigar[k] >> 4);
  return s->pos[i] + rl - 1;
}

static int64_t exotic_slot(const pl_read_soa* s, int64_t i) {
  int64_t lo = 0, hi = s->n_exotic;
  while (lo < hi) { int64_t m = (lo + hi) / 2; if (s->exo_read[m] < i) lo = m + 1; else hi = m; }
  return (lo < s->n_exotic && s->exo_read[lo] == i) ? lo : -1;
}

static int decode_read(Decoder* d, int64_t i, const pl_config* cfg, ReadView* r) {
  const pl_read_soa* s = d->s;
  int64_t o = s->seq_off[i];
  int32_t L = (int32_t)(s->seq_off[i + 1] - o);
  if (L + 1 > d->cap) {
    d->cap = 2 * (L + 1);
    d->bases = (uint8_t*)realloc(d->bases, (size_t)d->cap);
    d->quals = (int8_t*)realloc(d->quals, (size_t)d->cap);
  }
  if (s->flag[i] & PL_RF_EXOTIC) {
    int64_t slot = exotic_slot(s, i);
    if (slot < 0) return -1;
    int64_t nb = s->exo_off[slot];
    for (int32_t k = 0; k < L; k++) {
      int64_t q = nb + k;
      uint8_t byte = s->nib[q >> 1];
      d->bases[k] = (uint8_t)NIB[(q & 1) ? (byte & 0xf) : (byte >> 4)];
    }
  } else {
    for (int32_t k = 0; k < L; k++) {
      int64_t q = o + k;
      d->bases[k] = (uint8_t)"ACGT"[(s->bases2[q >> 2] >> (2 * (q & 3))) & 3];
    }
  }
  for (int32_t k = 0; k < L; k++)
    d->quals[k] = (s->flag[i] & PL_RF_NO_QUAL) ? (int8_t)cfg->default_qual : s->quals[o + k];
  r->length = L; r->bases = d->bases; r->quals = d->quals;
  r->mq = s->mapq[i]; r->flag = s->flag[i]; r->tlen = s->tlen[i];
  r->aStart = s->pos[i]; r->aEnd = alignment_end(s, i);
  r->cigar = s->cigar + s->cigar_off[i]; r->ncigar = s->cigar_off[i + 1] - s->cigar_off[i];
  r->index = i;
  /* CIGAR and SEQ must agree (htsjdk accepts such records silently, Pilon
   * then indexes past the arrays); a read without SEQ is allowed. */
  if (L > 0) {
    int64_t rl = 0;
    for (int64_t k = 0; k < r->ncigar; k++)
      if (consumesRead((int)(r->cigar[k] & 0xf))) rl += (int64_t)(r->cigar[k] >> 4);
    if (rl != L) return -2;
  }
  for (int64_t k = 0; k < r->ncigar; k++) if ((r->cigar[k] & 0xf) > 8) return -2;
  return 0;
}

/* BamFile.validateRead (BamFile.scala:101-105) */
static int validateRead(uint16_t flag, const pl_config* cfg) {
  return (cfg->nonpf || !(flag & 0x200)) && (cfg->duplicates || !(flag & 0x400)) && !(flag & 0x100);
}

/* ------------------------------------------------------------ fixes */
typedef struct { int32_t loc; const uint8_t* was; int32_t wasLen; const uint8_t* patch; int32_t patchLen; int64_t seq; } Fix;
typedef struct { Fix* v; int32_t n, cap; } FixList;

static void fix_prepend(FixList* l, Fix f) {               /* list ::= fix (kept reversed, fixed later) */
  if (l->n == l->cap) { l->cap = l->cap ? 2 * l->cap : 16; l->v = (Fix*)realloc(l->v, sizeof(Fix) * (size_t)l->cap); }
  l->v[l->n++] = f;
}
/* materialise a prepend-built list in List order (last prepended first) */
static void fix_as_list(const FixList* l, Fix* out) { for (int32_t i = 0; i < l->n; i++) out[i] = l->v[l->n - 1 - i]; }

static int cmp_fix_stable(const void* a, const void* b) {
  const Fix* x = (const Fix*)a; const Fix* y = (const Fix*)b;
  if (x->loc != y->loc) return x->loc < y->loc ? -1 : 1;
  return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

/* GenomeRegion.fixFixList (GenomeRegion.scala:557-593).  in[] is in List order;
 * returns the number of kept fixes written to out[] in coordinate order. */
static int32_t fixFixList(const Fix* in, int32_t n, Fix* out) {
  Fix* f = (Fix*)malloc(sizeof(Fix) * (size_t)(n ? n : 1));
  for (int32_t i = 0; i < n; i++) { f[i] = in[i]; f[i].seq = i; }
  qsort(f, (size_t)n, sizeof(Fix), cmp_fix_stable);       /* sortWith is a stable sort */
  int32_t m = 0, i = 0;
  if (n == 0) { free(f); return 0; }
  Fix cur = f[0]; i = 1;
  while (1) {
    if (i >= n) { out[m++] = cur; break; }
    Fix nx = f[i];
    int32_t s1 = cur.loc, e1 = cur.loc + ((cur.wasLen - 1) > 0 ? (cur.wasLen - 1) : 0);
    int32_t s2 = nx.loc, e2 = nx.loc + ((nx.wasLen - 1) > 0 ? (nx.wasLen - 1) : 0);
    if (s2 <= e1 && e2 >= s1) {                                       /* Region.overlaps */
      int32_t l1 = cur.wasLen + cur.patchLen, l2 = nx.wasLen + nx.patchLen;
      if (l1 < l2) cur = nx;
      i++;
    } else {
      out[m++] = cur; cur = nx; i++;
    }
  }
  free(f);
  return m;
}

/* GenomeRegion.fixIssues (GenomeRegion.scala:595-621), on a growable buffer. */
static void fixIssues(uint8_t** bases, int64_t* blen, const PRegion* R, const Fix* in, int32_t n) {
  Fix* kept = (Fix*)malloc(sizeof(Fix) * (size_t)(n ? n : 1));
  int32_t m = fixFixList(in, n, kept);
  for (int32_t k = m - 1; k >= 0; k--) {
    const Fix* f = &kept[k];
    int64_t start = f->loc - R->start;
    if (f->wasLen == f->patchLen) {
      for (int32_t i = 0; i < f->wasLen; i++) (*bases)[start + i] = f->patch[i];
    } else {
      int64_t nl = *blen - f->wasLen + f->patchLen;
      uint8_t* nb = (uint8_t*)malloc((size_t)(nl ? nl : 1));
      memcpy(nb, *bases, (size_t)start);
      memcpy(nb + start, f->patch, (size_t)f->patchLen);
      memcpy(nb + start + f->patchLen, *bases + start + f->wasLen, (size_t)(*blen - start - f->wasLen));
      free(*bases); *bases = nb; *blen = nl;
    }
  }
  free(kept);
}

/* ----------------------------------------------------------- results */
typedef struct orc_result {
  int32_t status;            /* pl_status */
  int64_t edge_read;
  char    edge_why[64];
  int32_t size;
  int32_t* pile;             /* [PL_NFIELDS*size] Int view (BaseSum Long truncated) */
  int64_t* bcq;              /* [8*size] exact BaseSum Longs                       */
  pl_call* calls;            /* [size] BaseCall on the final PileUp + dispositions */
  int32_t* coverage;         /* GenomeRegion.coverage etc. (GenomeRegion.scala:247-253) */
  int32_t* badCoverage;
  int32_t* physCoverage;
  int32_t* insertSize;
  int8_t*  weightedQual;
  int8_t*  weightedMq;
  int16_t* clips;
  pl_region_stats stats;
  char*    changes_text; int64_t changes_len;
  char*    vcf_text; int64_t vcf_len;   /* Vcf.writeRecord lines of the region (when asked) */
  char*    bed_text; int64_t bed_len;   /* the region's Pilon.bed lines (when asked) */
  int32_t* tracks;           /* [ORC_NTRACKS*size] Tracks.standardTracks values (when asked) */
  uint8_t* bases; int64_t bases_len;
  /* the called indel string at each INS/DEL change, packed: for change c,
   * indel_pos[c] (region index) and indel_str[indel_off[c] .. indel_off[c+1]) */
  int32_t  n_indel; int32_t* indel_pos; int64_t* indel_off; uint8_t* indel_str;
} orc_result;

void orc_result_free(orc_result* r) {
  if (!r) return;
  free(r->pile); free(r->bcq); free(r->calls); free(r->coverage); free(r->badCoverage);
  free(r->physCoverage); free(r->insertSize); free(r->weightedQual); free(r->weightedMq);
  free(r->clips); free(r->changes_text); free(r->vcf_text); free(r->bed_text); free(r->tracks); free(r->bases);
  free(r->indel_pos); free(r->indel_off); free(r->indel_str);
  free(r);
}

static void sb_append(char** buf, int64_t* len, int64_t* cap, const char* s, int64_t n) {
  if (*len + n + 1 > *cap) { *cap = 2 * (*len + n + 1) + 256; *buf = (char*)realloc(*buf, (size_t)*cap); }
  memcpy(*buf + *len, s, (size_t)n); *len += n; (*buf)[*len] = 0;
}
/* Region.regionString (Region.scala:42) */
static void region_string(char** buf, int64_t* len, int64_t* cap, const char* name, int64_t s, int64_t e) {
  char tmp[64];
  sb_append(buf, len, cap, name, (int64_t)strlen(name));
  int n = (e + 1 - s < 2) ? snprintf(tmp, sizeof tmp, ":%lld", (long long)s)
                          : snprintf(tmp, sizeof tmp, ":%lld-%lld", (long long)s, (long long)e);
  sb_append(buf, len, cap, tmp, n);
}

enum { K_SNP = 1, K_INS, K_DEL, K_AMB };

static uint8_t upper(uint8_t c) { return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; }

/* Bases.toIUPAC (Bases.scala:67-88) */
static uint8_t toIUPAC(int b1, int b2) {
  static const char T[16] = {0, 'A', 'C', 'M', 'G', 'R', 'S', 'V', 'T', 'W', 'Y', 'H', 'K', 'D', 'B', 'N'};
  int bits = (1 << b1) | (1 << b2);
  return (uint8_t)T[bits];
}

/*
 * Run one GenomeRegion end to end: BamFile.process over each read set,
 * postProcess Pass 1, identifyAndFixIssues (snps/indels), writeChanges.
 * contig: raw bytes, name: contig name, new_name/offset: writeChanges args.
 */
/* ------------------------------------------------------------------ VCF */
/* Java's "%.2f".format of a Float widened to Double (Vcf.scala:147): the
 * Formatter rounds the decimal digits of the value half up.  A widened float
 * has at most 24 significant bits, so no digit string shorter than its exact
 * expansion lies on a tie that the exact value does not: rounding the exact
 * binary value half up gives the same digits. */
static void fmt_af(char* out, float af) {
  int64_t q = 0;
  if (af > 0.0f) {
    int e;
    double m = frexp((double)af, &e);                    /* af = m * 2^e, m in [0.5, 1) */
    uint64_t mi = (uint64_t)ldexp(m, 24);                /* exact: 24 significant bits */
    uint64_t t = mi * 100u;                              /* af * 100 = t * 2^(e-24) */
    int sh = 24 - e;
    if (sh <= 0) q = (int64_t)(t << -sh);
    else if (sh < 63) {
      q = (int64_t)(t >> sh);
      if ((t & ((1ull << sh) - 1)) >= (1ull << (sh - 1))) q++;
    }
  }
  sprintf(out, "%lld.%02lld", (long long)(q / 100), (long long)(q % 100));
}

typedef struct {
  const pl_config* cfg; const uint8_t* contig; const char* name; int32_t start;
  int32_t minDepth; int diploid, vcfqe;
  char* buf; int64_t len, cap;
} VcfOut;

static void vcf_puts(VcfOut* o, const char* s) { sb_append(&o->buf, &o->len, &o->cap, s, (int64_t)strlen(s)); }
static void vcf_putn(VcfOut* o, const uint8_t* s, int32_t n) { sb_append(&o->buf, &o->len, &o->cap, (const char*)s, n); }
static void vcf_putl(VcfOut* o, int64_t v) { char t[32]; sprintf(t, "%lld", (long long)v); vcf_puts(o, t); }

/* Vcf.writeRecord (Vcf.scala:74-163) for region index i */
static void vcf_record(VcfOut* o, const PileUp* p, int32_t i, int embedded, int indelOkArg) {
  BaseCall bc; baseCall(p, o->cfg, &bc);
  const int indelOk = indelOkArg && i > 0;
  const int32_t locus = o->start + i;
  const uint8_t base = bc.base == 4 ? 'N' : (uint8_t)"ACGT"[bc.base];
  const uint8_t altB = (uint8_t)"ACGT"[bc.alt];
  /* bcString = callString(indelOk) (PileUp.scala:169-173) */
  const uint8_t* bcs = &base; int32_t bcl = 1;
  if (indelOk && (bc.isIns || bc.isDel)) { bcs = bc.indelStr; bcl = bc.indelLen; }
  const int32_t baseDP = (int32_t)p->qs[bc.bI];        /* bc.baseSum.toInt */
  const int32_t altBaseDP = (int32_t)p->qs[bc.alt];
  const int64_t depth = pu_depth(p);
  int32_t loc = locus;
  uint8_t rBase;
  uint8_t refS[2], altS[2];
  const uint8_t *refStr, *altStr; int32_t refL, altL;  /* REF and the called allele (cBase) */
  uint8_t* tmp = NULL;
  const char* callType;
  int32_t refDP, altDP;
  if (indelOk && !embedded && bc.isDel) {
    loc -= 1;
    rBase = upper(o->contig[loc - 1]);
    callType = bc.homoIndel ? "1/1" : "0/1";
    int32_t pc = delPct(p);
    tmp = (uint8_t*)malloc((size_t)bcl + 1); tmp[0] = rBase; memcpy(tmp + 1, bcs, (size_t)bcl);
    refStr = tmp; refL = bcl + 1; altS[0] = rBase; altStr = altS; altL = 1;
    refDP = 100 - pc; altDP = pc;
  } else if (indelOk && !embedded && bc.isIns) {
    loc -= 1;
    rBase = upper(o->contig[loc - 1]);
    callType = bc.homoIndel ? "1/1" : "0/1";
    int32_t pc = insPct(p);
    tmp = (uint8_t*)malloc((size_t)bcl + 1); tmp[0] = rBase; memcpy(tmp + 1, bcs, (size_t)bcl);
    refS[0] = rBase; refStr = refS; refL = 1; altStr = tmp; altL = bcl + 1;
    refDP = 100 - pc; altDP = pc;
  } else if (bc.homo) {
    rBase = upper(o->contig[loc - 1]);
    refS[0] = rBase; refStr = refS; refL = 1; altS[0] = base; altStr = altS; altL = 1;
    if (rBase == base || (bcl == 1 && bcs[0] == 'N')) { callType = "0/0"; refDP = baseDP; altDP = altBaseDP; }
    else { callType = "1/1"; refDP = altBaseDP; altDP = baseDP; }
  } else {
    rBase = upper(o->contig[loc - 1]);
    refS[0] = rBase; refStr = refS; refL = 1;
    callType = "0/1";
    if (rBase == base) { altS[0] = altB; refDP = baseDP; altDP = altBaseDP; }
    else { altS[0] = base; refDP = altBaseDP; altDP = baseDP; }
    altStr = altS; altL = 1;
  }
  /* filters: prepended, so the last added comes first */
  const char* f[3]; int nf = 0;
  if (embedded) f[nf++] = "Del";
  if (!o->diploid && callType[0] == '0' && callType[2] == '1') f[nf++] = "Amb";
  if (depth < o->minDepth) f[nf++] = "LowCov";
  const int dotAlt = (altL == 1 && altStr[0] == 'N') || (altL == refL && memcmp(altStr, refStr, (size_t)refL) == 0);
  const int ac = callType[0] == '0' ? (callType[2] == '0' ? 0 : 1) : 2;
  const int32_t dsum = jadd(refDP, altDP);
  const float af = (dsum > 0 && !dotAlt) ? (float)altDP / (float)dsum : 0.0f;
  char afs[32]; fmt_af(afs, af);
  const int64_t count = pu_count(p);
  const int64_t qsum = p->qs[0] + p->qs[1] + p->qs[2] + p->qs[3];
  const int64_t meanQual = roundDivL(qsum, roundDivL((int64_t)p->mqSum * count, depth));   /* PileUp.scala:64-67 */
  const int64_t meanMq = roundDivL((int64_t)p->mqSum - depth, depth);                     /* :70-72 */
  vcf_puts(o, o->name); vcf_puts(o, "\t"); vcf_putl(o, loc); vcf_puts(o, "\t.\t");
  vcf_putn(o, refStr, refL); vcf_puts(o, "\t");
  if (dotAlt) vcf_puts(o, "."); else vcf_putn(o, altStr, altL);
  vcf_puts(o, "\t");
  if (indelOk && bc.isDel) vcf_puts(o, "."); else vcf_putl(o, bc.score);
  vcf_puts(o, "\t");
  if (!nf) vcf_puts(o, "PASS");
  for (int k = 0; k < nf; k++) { if (k) vcf_puts(o, ";"); vcf_puts(o, f[k]); }
  vcf_puts(o, "\tDP="); vcf_putl(o, embedded ? count : depth);
  vcf_puts(o, ";TD="); vcf_putl(o, depth + p->badPair);
  vcf_puts(o, ";BQ="); vcf_putl(o, meanQual);
  vcf_puts(o, ";MQ="); vcf_putl(o, meanMq);
  vcf_puts(o, ";QD="); vcf_putl(o, bc.q);
  vcf_puts(o, ";BC=");
  for (int b = 0; b < 4; b++) { if (b) vcf_puts(o, ","); vcf_putl(o, p->bc[b]); }
  if (o->vcfqe) {                                       /* BaseSum.toString (BaseSum.scala:73) */
    vcf_puts(o, ";QE=");
    for (int b = 0; b < 4; b++) { if (b) vcf_puts(o, ","); vcf_putl(o, p->qs[b]); }
  } else {                                              /* BaseSum.toStringPct (BaseSum.scala:68-71) */
    vcf_puts(o, ";QP=");
    for (int b = 0; b < 4; b++) { if (b) vcf_puts(o, ","); vcf_putl(o, qsum == 0 ? 0 : ldiv64(100 * p->qs[b] + qsum / 2, qsum)); }
  }
  vcf_puts(o, ";PC="); vcf_putl(o, p->physCov);
  vcf_puts(o, ";IC="); vcf_putl(o, p->insertions);
  vcf_puts(o, ";DC="); vcf_putl(o, p->deletions);
  vcf_puts(o, ";XC="); vcf_putl(o, p->clips);
  vcf_puts(o, ";AC="); vcf_putl(o, ac);
  vcf_puts(o, ";AF="); vcf_puts(o, afs);
  vcf_puts(o, "\tGT\t"); vcf_puts(o, callType); vcf_puts(o, "\n");
  free(tmp);
  if (indelOk && bc.indel && !embedded) vcf_record(o, p, i, bc.isDel && bc.homoIndel, 0);
}

/* --------------------------------------------------------------- tracks */
/* The per-locus tracks of Tracks.standardTracks (Tracks.scala:33-56), in
 * that order; the Pilon.bed issue track is not among them. */
#define ORC_NTRACKS 13

/* GenomeRegion.smooth (GenomeRegion.scala:190-210), Int arithmetic.  The
 * window sum keeps input(0) (the first subtraction drops input(1)), and
 * result(window - half) is never written before it fills the left edge. */
static void smooth_int(const int32_t* in, int32_t n, int32_t window, int32_t* out) {
  const int32_t half = window / 2;
  int32_t accum = 0;
  memset(out, 0, sizeof(int32_t) * (size_t)n);
  for (int32_t i = 0; i < n; i++) {
    accum = jadd(accum, in[i]);
    if (i > window) {
      accum = jsub(accum, in[i - window]);
      out[i - half] = jdiv(jadd(accum, half), window);
    }
  }
  if (n > window) {
    for (int32_t i = 0; i < window - half; i++) out[i] = out[window - half];
    for (int32_t i = n - half; i < n; i++) out[i] = out[n - half - 1];
  } else {
    for (int32_t i = 0; i < n; i++) out[i] = jdiv(accum, n);
  }
}

/* GenomeRegion.computeGc (GenomeRegion.scala:158-188): GC count of a
 * 100-base circular buffer over the whole contig, stored at the window's
 * center (and at the last half window's own loci).  The scan runs over
 * 0-based contig indices but tests and indexes them with the region's
 * 1-based bounds (inRegion, index), so start/stop here are the 1-based
 * numbers and gc[k] holds the count when center == start + k. */
static void compute_gc(const uint8_t* contig, int64_t contig_len, int32_t start, int32_t stop, int8_t* gc) {
  const int window = 100, halfWindow = (window + 1) / 2;
  uint8_t buf[100];
  int32_t count = 0;
  for (int i = 0; i < window; i++) { buf[i] = (i & 1) == 0; count += buf[i]; }
  for (int64_t locus = 0; locus < contig_len; locus++) {
    const int64_t center = locus >= halfWindow ? locus - halfWindow : locus;
    const int bi = (int)(locus % window);
    const uint8_t c = contig[locus];
    const uint8_t g = (c == 'G' || c == 'C') ? 1 : (c == 'A' || c == 'T') ? 0 : buf[bi];
    count += (int32_t)g - (int32_t)buf[bi];
    buf[bi] = g;
    if (center >= start && center <= stop) gc[center - start] = (int8_t)count;
    if (locus >= start && locus <= stop && locus >= contig_len - halfWindow) gc[locus - start] = (int8_t)count;
  }
}

/* ------------------------------------------------------------ Pilon.bed */
typedef struct { int32_t start, stop; } IssueRegion;      /* Region(name, start, stop), loci */
typedef struct { IssueRegion* v; int32_t n, cap; } RegionList;

static void rl_push(RegionList* l, int32_t a, int32_t b) {
  if (l->n == l->cap) { l->cap = 2 * l->cap + 16; l->v = (IssueRegion*)realloc(l->v, sizeof(IssueRegion) * (size_t)l->cap); }
  l->v[l->n].start = a; l->v[l->n].stop = b; l->n++;
}

/* GenomeRegion.summaryRegions (GenomeRegion.scala:740-763): runs of loci
 * passing the test, a run closing at the first failing index more than
 * `slop` past its last passing one; regions within 100 of either end of the
 * GenomeRegion are dropped (nearEdge, :691). */
static RegionList summary_regions(const uint8_t* test, int32_t size, int32_t start, int32_t stop, int32_t slop) {
  RegionList all = {0}, out = {0};
  int32_t first = -1, last = -1;
  for (int32_t i = 0; i < size; i++) {
    if (test[i]) {
      last = i;
      if (first < 0) first = i;
    } else if (last >= 0 && i > last + slop) {
      rl_push(&all, start + first, start + last);
      first = -1;
      last = -1;
    }
  }
  if (last >= 0) rl_push(&all, start + first, start + last);
  for (int32_t k = 0; k < all.n; k++)
    if (!(all.v[k].start - start < 100 || stop - all.v[k].stop < 100)) rl_push(&out, all.v[k].start, all.v[k].stop);
  free(all.v);
  return out;
}

/* Tracks.regionsToBed (Tracks.scala:179-183) */
static void bed_append(char** buf, int64_t* len, int64_t* cap, const char* name, const RegionList* l,
                       const char* label, const char* rgb) {
  char line[512];
  for (int32_t k = 0; k < l->n; k++) {
    int n = snprintf(line, sizeof line, "%s\t%d\t%d\t%s\t0\t+\t%d\t%d\t%s\n", name, l->v[k].start - 1, l->v[k].stop,
                     label, l->v[k].start - 1, l->v[k].stop, rgb);
    if (n >= (int)sizeof line) {                            /* a long contig name */
      char* big = (char*)malloc((size_t)n + 1);
      snprintf(big, (size_t)n + 1, "%s\t%d\t%d\t%s\t0\t+\t%d\t%d\t%s\n", name, l->v[k].start - 1, l->v[k].stop, label,
               l->v[k].start - 1, l->v[k].stop, rgb);
      sb_append(buf, len, cap, big, n);
      free(big);
    } else {
      sb_append(buf, len, cap, line, n);
    }
  }
}

/* The region's lines of Tracks.makeBedTrack (Tracks.scala:194-216): its
 * unconfirmed, change, low-coverage, collapsed-repeat, duplication, gap and
 * break regions (GenomeRegion.scala:696-738) in that order.  The
 * reassembly outcomes that follow them are not produced (no reassembly on
 * this path). */
static void bed_lines(const pl_config* cfg, const char* name, int32_t start, int32_t stop, const uint8_t* contig,
                      const int32_t* coverage, const int32_t* badCoverage, const int16_t* clips,
                      const uint8_t* confirmed, const uint8_t* kindAt, const uint8_t* deleted,
                      const int16_t* copyNumber, const int32_t* fragCov, double fragMean, char** text, int64_t* len) {
  const int32_t size = stop + 1 - start;
  uint8_t* t = (uint8_t*)calloc((size_t)size, 1);
  int64_t totalBad = 0, totalGood = 0;                      /* pctBadOverall (:99-113), Long */
  for (int32_t i = 0; i < size; i++) { totalBad += badCoverage[i]; totalGood += coverage[i]; }
  const int64_t pctBadOverall = roundDivL(100 * totalBad, totalGood + totalBad);
  int64_t cap = 0;
  *text = NULL;
  *len = 0;
  RegionList L[7];
  for (int32_t i = 0; i < size; i++) t[i] = !confirmed[i];
  L[0] = summary_regions(t, size, start, stop, 100);                                    /* unConfirmedRegions */
  for (int32_t i = 0; i < size; i++) t[i] = kindAt[i] == K_SNP || kindAt[i] == K_INS || kindAt[i] == K_DEL || deleted[i];
  L[1] = summary_regions(t, size, start, stop, 1);                                      /* changeRegions */
  for (int32_t i = 0; i < size; i++)                                                    /* lowCoverage (:700) */
    t[i] = coverage[i] < cfg->min_min_depth && upper(contig[(int64_t)start + i - 1]) != 'N';
  L[2] = summary_regions(t, size, start, stop, 100);
  for (int32_t i = 0; i < size; i++) t[i] = copyNumber[i] > 1;
  RegionList hcn = summary_regions(t, size, start, stop, 100);                          /* highCopyNumberRegions */
  L[3] = (RegionList){0};
  for (int32_t k = 0; k < hcn.n; k++) {                     /* possibleCollapsedRepeats (:724-729) */
    const int32_t a = hcn.v[k].start - start, b = hcn.v[k].stop - start;
    int32_t d[2];
    for (int e = 0; e < 2; e++) {
      const int32_t i = e ? b : a;
      const int32_t l = fragCov[i > 100 ? i - 100 : 0], r = fragCov[i + 100 < size - 1 ? i + 100 : size - 1];
      const int32_t x = jsub(l, r);
      d[e] = x < 0 ? jsub(0, x) : x;                        /* delta (:672-678) */
    }
    if ((double)d[0] / fragMean >= 0.5 && (double)d[1] / fragMean >= 0.5) rl_push(&L[3], hcn.v[k].start, hcn.v[k].stop);
  }
  free(hcn.v);
  RegionList dup = summary_regions(t, size, start, stop, 2000);                        /* duplicationEvents (:731-737) */
  L[4] = (RegionList){0};
  for (int32_t k = 0; k < dup.n; k++)
    if (dup.v[k].stop + 1 - dup.v[k].start > 10000) rl_push(&L[4], dup.v[k].start, dup.v[k].stop);
  free(dup.v);
  for (int32_t i = 0; i < size; i++) t[i] = upper(contig[(int64_t)start + i - 1]) == 'N';
  RegionList g = summary_regions(t, size, start, stop, 100);                            /* gaps (:698) */
  L[5] = (RegionList){0};
  for (int32_t k = 0; k < g.n; k++)
    if (g.v[k].stop + 1 - g.v[k].start >= 10) rl_push(&L[5], g.v[k].start, g.v[k].stop);
  free(g.v);
  for (int32_t i = 0; i < size; i++) {                      /* breakp (:719) */
    const int32_t cov = coverage[i], bad = badCoverage[i];
    const int lowCov = cov < cfg->min_min_depth && upper(contig[(int64_t)start + i - 1]) != 'N';
    const int highClip = cov >= cfg->min_min_depth && pctI(clips[i], cov) >= 33;
    const int tooBad = (int64_t)pctI(bad, jadd(cov, bad)) > pctBadOverall + 20;
    const int32_t l = fragCov[i > 100 ? i - 100 : 0], r = fragCov[i + 100 < size - 1 ? i + 100 : size - 1];
    const int32_t dip = jadd(jsub(l, fragCov[i]), jsub(r, fragCov[i]));
    t[i] = lowCov || highClip || tooBad || (double)dip / fragMean >= 1.5;
  }
  RegionList br = summary_regions(t, size, start, stop, 200);                           /* possibleBreaks (:720) */
  L[6] = (RegionList){0};
  for (int32_t k = 0; k < br.n; k++) {                      /* filter !nearAny(gaps, 300) (Region.scala:37-40) */
    int near = 0;
    for (int32_t q = 0; q < L[5].n && !near; q++) {
      const IssueRegion G = L[5].v[q], B = br.v[k];
      near = (B.start <= G.stop && B.stop >= G.start) || abs(B.stop - G.start) <= 300 || abs(B.start - G.stop) <= 300;
    }
    if (!near) rl_push(&L[6], br.v[k].start, br.v[k].stop);
  }
  free(br.v);
  static const char* labels[7] = {"?", "X", "LowCov", "Copy#", "Duplication", "Gap", "Break"};
  static const char* rgbs[7] = {"0,128,128", "255,0,0", "128,0,128", "128,128,0", "255,128,0", "0,0,0", "255,0,255"};
  for (int k = 0; k < 7; k++) {
    bed_append(text, len, &cap, name, &L[k], labels[k], rgbs[k]);
    free(L[k].v);
  }
  free(t);
}

/* set_types[si] (NULL: all fragment BAMs) is the BAM's type: 0 frags, 1
 * jumps (whose depth increments processBam leaves out of fragCoverage,
 * GenomeRegion.scala:287-300), 2 unpaired, 3 nanopore and 4 pacbio
 * (unpaired BAMs of longReadType 1 and 2, BamFile.scala:43-47). */
static int fix_eq(const Fix* a, const Fix* b) {            /* tuple equality of (loc, was, patch) */
  return a->loc == b->loc && a->wasLen == b->wasLen && a->patchLen == b->patchLen &&
         (a->wasLen == 0 || !memcmp(a->was, b->was, (size_t)a->wasLen)) &&
         (a->patchLen == 0 || !memcmp(a->patch, b->patch, (size_t)a->patchLen));
}

/* big_*: bigFixList (the local reassembly's fixes, GenomeRegion.scala:
 * 382-410) in List order, applied after the small fixes, listed in the
 * changes and written as Vcf.writeFixRecord records (n_big 0: none). */
orc_result* orc_region_run_big(const pl_config* cfg, const uint8_t* contig, int64_t contig_len,
                               const char* name, int32_t start, int32_t stop,
                               const pl_read_soa* const* sets, const int32_t* set_types, int32_t n_sets,
                               const char* new_name, int64_t offset, int32_t vcf_flags,
                               const int32_t* big_loc, const char* const* big_was, const char* const* big_patch,
                               int32_t n_big) {
  orc_result* res = (orc_result*)calloc(1, sizeof(orc_result));
  if (start < 1 || stop < start || stop > contig_len) { res->status = PL_E_INVAL; return res; }  /* GenomeRegion.scala:34 */
  PRegion R; memset(&R, 0, sizeof R);
  R.start = start; R.stop = stop; R.size = stop + 1 - start;
  R.pu = (PileUp*)calloc((size_t)R.size, sizeof(PileUp));
  const int32_t size = R.size;
  /* processBam's fragCoverage (GenomeRegion.scala:287-300): each non-jumps
   * BAM's depth increment, before Pass 1 mutates any deletion count */
  int32_t* fragCov = (int32_t*)calloc((size_t)size, 4);
  int32_t* covBefore = (int32_t*)calloc((size_t)size, 4);

  /* BamFile.process (BamFile.scala:108-148): query ±10 kb, validateRead, addRead */
  int64_t qs = (int64_t)start - 10000 > 0 ? (int64_t)start - 10000 : 0;
  int64_t qe = (int64_t)stop + 10000 < contig_len ? (int64_t)stop + 10000 : contig_len;
  for (int32_t si = 0; si < n_sets && !R.edge; si++) {
    const pl_read_soa* s = sets[si];
    const int jumps = set_types && set_types[si] == 1;
    const int32_t longRead = !set_types ? 0 : set_types[si] == 3 ? LR_NANOPORE : set_types[si] == 4 ? LR_PACBIO : 0;
    if (!jumps) for (int32_t i = 0; i < size; i++) covBefore[i] = (int32_t)pu_depth(&R.pu[i]);
    Decoder d = {s, NULL, NULL, 0};
    for (int64_t i = 0; i < s->n_reads; i++) {
      int32_t aEnd = alignment_end(s, i);
      int64_t spanEnd = aEnd != 0 ? aEnd : s->pos[i];
      if (!(s->pos[i] <= qe && spanEnd >= qs)) continue;           /* queryOverlapping */
      if (!validateRead(s->flag[i], cfg)) continue;
      ReadView r;
      int rc = decode_read(&d, i, cfg, &r);
      if (rc != 0) { res->status = PL_E_INVAL; res->edge_read = i; snprintf(res->edge_why, sizeof res->edge_why, "malformed read"); break; }
      /* aligned-base unit: M/=/X bases whose locus is in the region */
      int32_t ro = 0, fo = 0;
      for (int64_t k = 0; k < r.ncigar; k++) {
        int op = (int)(r.cigar[k] & 0xf); int32_t len = (int32_t)(r.cigar[k] >> 4);
        if (op == 0 || op == 7 || op == 8) {
          int64_t a = (int64_t)r.aStart + fo, b = a + len - 1;
          if (a < start) a = start;
          if (b > stop) b = stop;
          if (b >= a) R.alignedBases += b - a + 1;
        }
        if (consumesRead(op)) ro += len;
        if (consumesRef(op)) fo += len;
      }
      pr_addRead(&R, &r, contig, contig_len, longRead, cfg);
      if (R.edge) break;
    }
    free(d.bases); free(d.quals);
    if (!jumps)
      for (int32_t i = 0; i < size; i++) fragCov[i] = jadd(fragCov[i], jsub((int32_t)pu_depth(&R.pu[i]), covBefore[i]));
    if (res->status) break;
  }
  free(covBefore);
  if (R.edge) {
    res->status = PL_E_EDGE; res->edge_read = R.edgeRead;
    snprintf(res->edge_why, sizeof res->edge_why, "%s", R.edgeWhy);
  }
  if (res->status) {
    for (int32_t i = 0; i < size; i++) { free(R.pu[i].ins.v); free(R.pu[i].del.v); }
    free(R.pu); arena_free(&R.ar); free(fragCov);
    return res;
  }

  /* GenomeRegion.postProcess (GenomeRegion.scala:214-272) */
  pr_computePhysCov(&R);
  int64_t meanCoverage = roundDivL(R.baseCount, (int64_t)size);          /* PileUpRegion.coverage */
  int32_t minDepth;
  if (cfg->min_depth >= 1) minDepth = cfg->min_depth >= 2147483647.0 ? INT32_MAX : (int32_t)cfg->min_depth;  /* toInt saturates */
  else {
    int32_t md = (int32_t)orc_java_round(cfg->min_depth * (double)meanCoverage);
    minDepth = md > cfg->min_min_depth ? md : cfg->min_min_depth;
  }
  uint8_t* confirmed = (uint8_t*)calloc((size_t)size, 1);
  uint8_t* deleted = (uint8_t*)calloc((size_t)size, 1);
  uint8_t* kindAt = (uint8_t*)calloc((size_t)size, 1);                  /* changeMap */
  res->coverage = (int32_t*)calloc((size_t)size, 4);
  res->badCoverage = (int32_t*)calloc((size_t)size, 4);
  res->physCoverage = (int32_t*)calloc((size_t)size, 4);
  res->insertSize = (int32_t*)calloc((size_t)size, 4);
  res->weightedQual = (int8_t*)calloc((size_t)size, 1);
  res->weightedMq = (int8_t*)calloc((size_t)size, 1);
  res->clips = (int16_t*)calloc((size_t)size, 2);
  const int fixamb = cfg->iupac || cfg->fix_amb;
  if (R.readCount != 0) {
    for (int32_t i = 0; i < size; i++) {
      PileUp* pu = &R.pu[i];
      int64_t n = pu_depth(pu);
      BaseCall bc; baseCall(pu, cfg, &bc);
      uint8_t b = bc.base == 4 ? 'N' : (uint8_t)"ACGT"[bc.base];
      uint8_t rr = upper(contig[(int64_t)start + i - 1]);              /* refBase (GenomeRegion.scala:783-787) */
      res->coverage[i] = (int32_t)n;
      res->badCoverage[i] = pu->badPair;
      res->physCoverage[i] = pu->physCov;
      res->insertSize[i] = pu->insertSize;
      int64_t qsum = pu->qs[0] + pu->qs[1] + pu->qs[2] + pu->qs[3];
      res->weightedQual[i] = (int8_t)roundDivL(qsum, (int64_t)pu->mqSum);   /* PileUp.scala:60-62 */
      res->weightedMq[i] = (int8_t)roundDivL(qsum, (int64_t)pu->qSum);     /* PileUp.scala:56-58 */
      res->clips[i] = (int16_t)pu->clips;
      if (n >= minDepth && rr != 'N' && !deleted[i] && bc.called) {
        if (bc.homo && b == rr && bc.highConf && !bc.indel) confirmed[i] = 1;
        else if (bc.isIns && bc.homoIndel) kindAt[i] = K_INS;
        else if (bc.isDel && bc.homoIndel) {
          kindAt[i] = K_DEL;
          for (int32_t j = 1; j < bc.indelLen; j++) {
            deleted[i + j] = 1;
            R.pu[i + j].deletions = jadd(R.pu[i + j].deletions, pu->deletions);
          }
        } else if (b != rr && bc.score > 0) {
          if (bc.homo) kindAt[i] = K_SNP;
          else if (fixamb || (uint8_t)"ACGT"[bc.alt] != rr) kindAt[i] = K_AMB;
        }
      }
    }
  }

  /* Pass 2 (GenomeRegion.scala:273-283): copy number, for the tracks
     (Tracks.scala:33-190) and the VCF's duplication records */
  int16_t* copyNumber = NULL;
  if (vcf_flags & (1 | 8 | 16)) {
    copyNumber = (int16_t*)calloc((size_t)size, 2);
    if (R.readCount != 0) {                                     /* Pass 2 runs after the early return */
      double sum = 0.0;
      for (int32_t i = 0; i < size; i++) sum += (double)fragCov[i];   /* NormalDistribution.mean */
      const double baseCov = sum / (double)size;
      int32_t* sm = (int32_t*)malloc(sizeof(int32_t) * (size_t)size);
      smooth_int(fragCov, size, 200, sm);
      for (int32_t i = 0; i < size; i++)
        copyNumber[i] = baseCov > 0 ? (int16_t)orc_java_round((double)sm[i] / baseCov) : 0;
      free(sm);
    }
  }
  if (vcf_flags & (8 | 16)) {
    int32_t* t = res->tracks = (int32_t*)calloc((size_t)ORC_NTRACKS * (size_t)size, 4);
    int8_t* gc = (int8_t*)calloc((size_t)size, 1);
    compute_gc(contig, contig_len, start, stop, gc);
    for (int32_t i = 0; i < size; i++) {
      const int32_t cov = res->coverage[i], bad = res->badCoverage[i];
      const int32_t l = fragCov[i > 100 ? i - 100 : 0], r = fragCov[i + 100 < size - 1 ? i + 100 : size - 1];
      const int32_t c = fragCov[i];
      const int32_t d = jsub(l, r);
      t[0 * (size_t)size + i] = (kindAt[i] == K_SNP || kindAt[i] == K_INS || kindAt[i] == K_DEL);   /* changed */
      t[1 * (size_t)size + i] = confirmed[i] ? 0 : 1;
      t[2 * (size_t)size + i] = (int32_t)copyNumber[i] - 1;
      t[3 * (size_t)size + i] = cov;
      t[4 * (size_t)size + i] = bad;
      t[5 * (size_t)size + i] = jadd(cov, bad) > 0 ? jdiv(jmul(bad, 100), jadd(cov, bad)) : 0;
      t[6 * (size_t)size + i] = d < 0 ? jsub(0, d) : d;                                        /* delta */
      t[7 * (size_t)size + i] = jadd(jsub(l, c), jsub(r, c));                                  /* dip */
      t[8 * (size_t)size + i] = res->physCoverage[i];
      t[9 * (size_t)size + i] = res->clips[i];
      t[10 * (size_t)size + i] = res->weightedQual[i];
      t[11 * (size_t)size + i] = res->weightedMq[i];
      t[12 * (size_t)size + i] = gc[i];
    }
    free(gc);
  }

  /* final PileUp state and BaseCall per position */
  res->size = size;
  res->pile = (int32_t*)malloc(sizeof(int32_t) * (size_t)size * PL_NFIELDS);
  res->bcq = (int64_t*)malloc(sizeof(int64_t) * (size_t)size * 8);
  res->calls = (pl_call*)calloc((size_t)size, sizeof(pl_call));
  int nConfirmed = 0;
  for (int32_t i = 0; i < size; i++) {
    PileUp* p = &R.pu[i];
    for (int b = 0; b < 4; b++) {
      res->bcq[b * (size_t)size + i] = p->bc[b];
      res->bcq[(4 + b) * (size_t)size + i] = p->qs[b];
      res->pile[(PL_F_COUNT_A + b) * (size_t)size + i] = (int32_t)p->bc[b];
      res->pile[(PL_F_QUAL_A + b) * (size_t)size + i] = (int32_t)p->qs[b];
    }
    res->pile[PL_F_MQSUM * (size_t)size + i] = p->mqSum;
    res->pile[PL_F_QSUM * (size_t)size + i] = p->qSum;
    res->pile[PL_F_PHYSCOV * (size_t)size + i] = p->physCov;
    res->pile[PL_F_INSERTSIZE * (size_t)size + i] = p->insertSize;
    res->pile[PL_F_BADPAIR * (size_t)size + i] = p->badPair;
    res->pile[PL_F_DELETIONS * (size_t)size + i] = p->deletions;
    res->pile[PL_F_DELQUAL * (size_t)size + i] = p->delQual;
    res->pile[PL_F_INSERTIONS * (size_t)size + i] = p->insertions;
    res->pile[PL_F_INSQUAL * (size_t)size + i] = p->insQual;
    res->pile[PL_F_CLIPS * (size_t)size + i] = p->clips;
    BaseCall bc; baseCall(p, cfg, &bc);
    pl_call* c = &res->calls[i];
    c->base = (uint8_t)bc.base; c->alt = (uint8_t)bc.alt; c->q = (int32_t)bc.q;
    uint16_t f = 0;
    if (bc.homo) f |= PL_CF_HOMO;
    if (bc.highConf) f |= PL_CF_HIGHCONF;
    if (bc.called) f |= PL_CF_CALLED;
    if (bc.indel) f |= PL_CF_INDEL;
    if (bc.homoIndel) f |= PL_CF_HOMOINDEL;
    if (bc.isIns) f |= PL_CF_INSERTION;
    if (bc.isDel) f |= PL_CF_DELETION;
    if (bc.score > 0) f |= PL_CF_SCOREPOS;
    int disp = PL_D_NONE;
    if (confirmed[i]) disp = PL_D_CONFIRMED;
    else if (kindAt[i] == K_SNP) disp = PL_D_SNP;
    else if (kindAt[i] == K_INS) disp = PL_D_INS;
    else if (kindAt[i] == K_DEL) disp = PL_D_DEL;
    else if (kindAt[i] == K_AMB) disp = PL_D_AMB;
    f |= (uint16_t)(disp << PL_CF_DISP_SHIFT);
    if (deleted[i]) f |= PL_CF_DELETED;
    /* the output format's mark of a Long qualSum beyond int32 (the pileup
       planes keep its low 32 bits; bcq holds the exact value) */
    for (int b = 0; b < 4; b++) if (p->qs[b] != (int64_t)(int32_t)p->qs[b]) f |= PL_CF_WIDE;
    c->flags = f;
    nConfirmed += confirmed[i];
  }

  /* excludeMotifs (GenomeRegion.scala:219,432-439), when every BAM is a
   * long-read or other non-frags BAM of a long-read run (longReadOnly,
   * :147): loci starting a homopolymer of 4 within the region, or, with a
   * nanopore BAM, the middle of CC_GG (upper-cased region bases), take no fix */
  int longReadOnly = cfg->longread != 0, nano = 0;
  for (int32_t si = 0; si < n_sets; si++) {
    const int32_t t = set_types ? set_types[si] : 0;
    if (t == 0) longReadOnly = 0;
    if (t == 3) nano = 1;
  }

  /* identifyAndFixIssues, snps/indels path (GenomeRegion.scala:307-380,413) */
  FixList snpL = {0}, smallL = {0};
  int snps = 0, ins = 0, dels = 0, insBases = 0, delBases = 0, amb = 0;
  res->indel_pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)(size + 1));
  res->indel_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(size + 2));
  int64_t strCap = 1024, strLen = 0; res->indel_str = (uint8_t*)malloc((size_t)strCap);
  res->indel_off[0] = 0;
  static const uint8_t dot = '.';
  (void)dot;
  uint8_t* oneChar = (uint8_t*)malloc((size_t)size * 2 + 2);             /* storage for 1-char strings */
  for (int32_t i = 0; i < size; i++) {                                   /* changeList is sorted */
    if (!kindAt[i]) continue;
    int32_t loc = start + i;
    uint8_t rBase = upper(contig[(int64_t)loc - 1]);
    BaseCall bc; baseCall(&R.pu[i], cfg, &bc);
    uint8_t cBase = bc.base == 4 ? 'N' : (uint8_t)"ACGT"[bc.base];
    oneChar[2 * i] = rBase; oneChar[2 * i + 1] = cBase;
    Fix f; memset(&f, 0, sizeof f); f.loc = loc;
    int excluded = 0;
    if (longReadOnly) {
      int32_t run = 1;                                                   /* homoRun(loc) (:416-422) */
      while (loc + run <= stop && upper(contig[(int64_t)loc + run - 1]) == rBase) run++;
      excluded = run >= 4 ||
                 (nano && loc - 2 >= start && loc + 2 <= stop &&           /* nanoporeExclude (:424-430) */
                  upper(contig[loc - 3]) == 'C' && upper(contig[loc - 2]) == 'C' &&
                  upper(contig[loc]) == 'G' && upper(contig[loc + 1]) == 'G');
    }
    if (!excluded) switch (kindAt[i]) {
      case K_SNP:
        if (cfg->fix_snps) { f.was = &oneChar[2 * i]; f.wasLen = 1; f.patch = &oneChar[2 * i + 1]; f.patchLen = 1; fix_prepend(&snpL, f); }
        snps++;
        break;
      case K_AMB:
        if (cfg->fix_snps && !cfg->longread) {
          if (cfg->iupac) {
            oneChar[2 * i + 1] = toIUPAC(bc.base == 4 ? 0 : bc.base, bc.alt);
            f.was = &oneChar[2 * i]; f.wasLen = 1; f.patch = &oneChar[2 * i + 1]; f.patchLen = 1; fix_prepend(&smallL, f);
          } else {
            f.was = &oneChar[2 * i]; f.wasLen = 1; f.patch = &oneChar[2 * i + 1]; f.patchLen = 1; fix_prepend(&snpL, f);
          }
          amb++;
        }
        break;
      case K_INS:
        if (cfg->fix_indels) { f.was = NULL; f.wasLen = 0; f.patch = bc.indelStr; f.patchLen = bc.indelLen; fix_prepend(&smallL, f); }
        ins++; insBases += bc.indelLen;
        break;
      case K_DEL:
        if (cfg->fix_indels) { f.was = bc.indelStr; f.wasLen = bc.indelLen; f.patch = NULL; f.patchLen = 0; fix_prepend(&smallL, f); }
        dels++; delBases += bc.indelLen;
        break;
    }
    if (kindAt[i] == K_INS || kindAt[i] == K_DEL) {
      if (strLen + bc.indelLen > strCap) { strCap = 2 * (strLen + bc.indelLen); res->indel_str = (uint8_t*)realloc(res->indel_str, (size_t)strCap); }
      memcpy(res->indel_str + strLen, bc.indelStr, (size_t)bc.indelLen);
      strLen += bc.indelLen;
      res->indel_pos[res->n_indel] = i;
      res->indel_off[++res->n_indel] = strLen;
    }
  }

  /* region bases, then fixIssues(snpFixList); fixIssues(smallFixList ++ bigFixList) */
  int64_t blen = size;
  uint8_t* bases = (uint8_t*)malloc((size_t)size);
  memcpy(bases, contig + start - 1, (size_t)size);                       /* GenomeRegion.refBases */
  Fix* tmp = (Fix*)malloc(sizeof(Fix) * (size_t)(snpL.n + smallL.n + n_big + 1));
  Fix* bigV = (Fix*)calloc((size_t)n_big + 1, sizeof(Fix));
  for (int32_t k = 0; k < n_big; k++) {
    bigV[k].loc = big_loc[k];
    bigV[k].was = (const uint8_t*)big_was[k]; bigV[k].wasLen = (int32_t)strlen(big_was[k]);
    bigV[k].patch = (const uint8_t*)big_patch[k]; bigV[k].patchLen = (int32_t)strlen(big_patch[k]);
  }
  fix_as_list(&snpL, tmp);
  fixIssues(&bases, &blen, &R, tmp, snpL.n);
  fix_as_list(&smallL, tmp);
  memcpy(tmp + smallL.n, bigV, sizeof(Fix) * (size_t)n_big);
  fixIssues(&bases, &blen, &R, tmp, smallL.n + n_big);

  /* writeChanges (GenomeRegion.scala:645-657) over snpFixList ++
   * smallFixList ++ bigFixList */
  fix_as_list(&snpL, tmp);
  fix_as_list(&smallL, tmp + snpL.n);
  memcpy(tmp + snpL.n + smallL.n, bigV, sizeof(Fix) * (size_t)n_big);
  Fix* kept = (Fix*)malloc(sizeof(Fix) * (size_t)(snpL.n + smallL.n + n_big + 1));
  int32_t nk = fixFixList(tmp, snpL.n + smallL.n + n_big, kept);
  char* txt = NULL; int64_t tl = 0, tc = 0;
  sb_append(&txt, &tl, &tc, "", 0);
  int64_t delta = 0;
  for (int32_t k = 0; k < nk; k++) {
    const Fix* f = &kept[k];
    int64_t newLoc = f->loc + delta;
    region_string(&txt, &tl, &tc, name, f->loc, (int64_t)f->loc + f->wasLen - 1);
    sb_append(&txt, &tl, &tc, " ", 1);
    region_string(&txt, &tl, &tc, new_name, newLoc + offset, newLoc + offset + f->patchLen - 1);
    sb_append(&txt, &tl, &tc, " ", 1);
    if (f->wasLen == 0) sb_append(&txt, &tl, &tc, ".", 1); else sb_append(&txt, &tl, &tc, (const char*)f->was, f->wasLen);
    sb_append(&txt, &tl, &tc, " ", 1);
    if (f->patchLen == 0) sb_append(&txt, &tl, &tc, ".", 1); else sb_append(&txt, &tl, &tc, (const char*)f->patch, f->patchLen);
    sb_append(&txt, &tl, &tc, "\n", 1);
    delta += f->patchLen - f->wasLen;
  }
  res->changes_text = txt; res->changes_len = tl;
  res->bases = bases; res->bases_len = blen;

  res->stats.read_count = R.readCount;
  res->stats.base_count = R.baseCount;
  res->stats.coverage = meanCoverage;
  res->stats.min_depth = minDepth;
  res->stats.n_confirmed = nConfirmed;
  res->stats.snps = snps; res->stats.amb = amb; res->stats.ins = ins; res->stats.ins_bases = insBases;
  res->stats.dels = dels; res->stats.del_bases = delBases;
  res->stats.n_unknown_ops = (int32_t)R.unknownOps;
  res->stats.aligned_bases = R.alignedBases;

  /* GenomeRegion.writeVcf (GenomeRegion.scala:623-643): every locus,
   * embedded when deleted, each duplicationEvents region (copy number > 1
   * with slop 2000, longer than 10,000, :731-737) announced by Vcf.writeDup
   * (Vcf.scala:194-202) before the records of its first locus, then at the
   * locus of each kept fix from bigFixList and not from smallFixList its
   * Vcf.writeFixRecord record (Vcf.scala:178-192), the loci it replaces
   * marked deleted from there on */
  if (vcf_flags & 1) {
    VcfOut o; memset(&o, 0, sizeof o);
    o.cfg = cfg; o.contig = contig; o.name = name; o.start = start; o.minDepth = minDepth;
    o.diploid = (vcf_flags & 2) != 0; o.vcfqe = (vcf_flags & 4) != 0;
    sb_append(&o.buf, &o.len, &o.cap, "", 0);
    uint8_t* t = (uint8_t*)calloc((size_t)size, 1);
    for (int32_t i = 0; i < size; i++) t[i] = copyNumber[i] > 1;
    RegionList dup = summary_regions(t, size, start, stop, 2000);
    free(t);
    int32_t d = 0, fk = 0;
    for (int32_t i = 0; i < size && !res->status; i++) {
      while (d < dup.n && dup.v[d].stop + 1 - dup.v[d].start <= 10000) d++;
      if (d < dup.n && dup.v[d].start == start + i) {
        const int32_t loc = dup.v[d].start - 1;                 /* writeDup: refBase(dup.start - 1) */
        vcf_puts(&o, name); vcf_puts(&o, "\t"); vcf_putl(&o, loc); vcf_puts(&o, "\t.\t");
        const uint8_t rb = upper(contig[loc - 1]);
        vcf_putn(&o, &rb, 1);
        vcf_puts(&o, "\t<DUP>\t.\tPASS\tSVTYPE=DUP;SVLEN="); vcf_putl(&o, dup.v[d].stop + 1 - dup.v[d].start);
        vcf_puts(&o, ";END="); vcf_putl(&o, dup.v[d].stop); vcf_puts(&o, ";IMPRECISE\tGT\t./.\n");
        d++;
      }
      if (fk < nk && kept[fk].loc == start + i) {
        const Fix* f = &kept[fk];
        int inBig = 0, inSmall = 0;
        for (int32_t k = 0; k < n_big && !inBig; k++) inBig = fix_eq(f, &bigV[k]);
        for (int32_t k = 0; k < smallL.n && !inSmall; k++) inSmall = fix_eq(f, &smallL.v[k]);
        if (inBig && !inSmall) {
          const int32_t loc = f->loc - 1;
          if (loc < start) { res->status = PL_E_EDGE; break; }  /* refBase's require(inRegion(loc)) */
          const uint8_t rb = upper(contig[loc - 1]);
          const int64_t refLen = 1 + f->wasLen, altLen = 1 + f->patchLen, svlen = altLen - refLen;
          vcf_puts(&o, name); vcf_puts(&o, "\t"); vcf_putl(&o, loc); vcf_puts(&o, "\t.\t");
          vcf_putn(&o, &rb, 1); vcf_putn(&o, f->was, f->wasLen); vcf_puts(&o, "\t");
          vcf_putn(&o, &rb, 1); vcf_putn(&o, f->patch, f->patchLen);
          vcf_puts(&o, "\t.\tPASS\tSVTYPE="); vcf_puts(&o, svlen < 0 ? "DEL" : "INS");
          vcf_puts(&o, ";SVLEN="); vcf_putl(&o, svlen); vcf_puts(&o, ";END="); vcf_putl(&o, loc + refLen - 1);
          if (f->patchLen && memchr(f->patch, 'N', (size_t)f->patchLen)) vcf_puts(&o, ";IMPRECISE");
          vcf_puts(&o, "\tGT\t1/1\n");
          for (int32_t j = 0; j < f->wasLen; j++) deleted[i + j] = 1;
        }
        fk++;
      }
      vcf_record(&o, &R.pu[i], i, deleted[i], 1);
    }
    free(dup.v);
    res->vcf_text = o.buf; res->vcf_len = o.len;
  }
  /* Pilon.bed comes after the VCF (Pilon.scala:111-116), so its change
   * regions see the loci writeVcf marked deleted */
  if ((vcf_flags & 16) && !res->status) {
    double fsum = 0.0;
    for (int32_t i = 0; i < size; i++) fsum += (double)fragCov[i];
    bed_lines(cfg, name, start, stop, contig, res->coverage, res->badCoverage, res->clips, confirmed, kindAt,
              deleted, copyNumber, fragCov, fsum / (double)size, &res->bed_text, &res->bed_len);
  }
  free(copyNumber);
  free(bigV);

  free(tmp); free(kept); free(snpL.v); free(smallL.v); free(oneChar);
  free(confirmed); free(deleted); free(kindAt); free(fragCov);
  for (int32_t i = 0; i < size; i++) { free(R.pu[i].ins.v); free(R.pu[i].del.v); }
  free(R.pu);
  /* indel strings point into the arena: copied above, so it can go */
  arena_free(&R.ar);
  return res;
}

orc_result* orc_region_run_typed(const pl_config* cfg, const uint8_t* contig, int64_t contig_len,
                                 const char* name, int32_t start, int32_t stop,
                                 const pl_read_soa* const* sets, const int32_t* set_types, int32_t n_sets,
                                 const char* new_name, int64_t offset, int32_t vcf_flags) {
  return orc_region_run_big(cfg, contig, contig_len, name, start, stop, sets, set_types, n_sets, new_name, offset,
                            vcf_flags, NULL, NULL, NULL, 0);
}

orc_result* orc_region_run(const pl_config* cfg, const uint8_t* contig, int64_t contig_len,
                           const char* name, int32_t start, int32_t stop,
                           const pl_read_soa* const* sets, int32_t n_sets,
                           const char* new_name, int64_t offset, int32_t vcf_flags) {
  return orc_region_run_typed(cfg, contig, contig_len, name, start, stop, sets, NULL, n_sets, new_name, offset,
                              vcf_flags);
}

/* Exposed for known-answer tests of the small helpers. */
void orc_smooth(const int32_t* in, int32_t n, int32_t window, int32_t* out) { smooth_int(in, n, window, out); }
int32_t orc_summary_regions(const uint8_t* test, int32_t size, int32_t start, int32_t slop, int32_t* out, int32_t cap) {
  RegionList l = summary_regions(test, size, start, start + size - 1, slop);
  const int32_t n = l.n;
  for (int32_t k = 0; k < n && k < cap; k++) { out[2 * k] = l.v[k].start; out[2 * k + 1] = l.v[k].stop; }
  free(l.v);
  return n;
}
void orc_gc(const uint8_t* contig, int64_t len, int32_t start, int32_t stop, int8_t* gc) { compute_gc(contig, len, start, stop, gc); }
void orc_fmt_af(float af, char* out) { fmt_af(out, af); }
int32_t orc_roundDivI(int32_t n, int32_t d) { return roundDivI(n, d); }
int64_t orc_roundDivL(int64_t n, int64_t d) { return roundDivL(n, d); }

/* BaseCall on a hand-made PileUp: counts bc[4], qs[4], and the Int fields in
 * PileUp order (mqSum,qSum,deletions,insertions,insQual,delQual); indel lists
 * are given as packed strings (ins then del).  Used to pin BaseCall and
 * hetIndelCall on exhaustive small cases. */
void orc_basecall(const pl_config* cfg, const int64_t* bc4, const int64_t* qs4,
                  int32_t mqSum, int32_t qSum, int32_t deletions, int32_t insertions,
                  int32_t insQual, int32_t delQual,
                  int32_t n_ins, const int32_t* ins_len, const uint8_t* ins_bytes,
                  int32_t n_del, const int32_t* del_len, const uint8_t* del_bytes,
                  pl_call* out, int32_t* indel_len, int64_t* score) {
  PileUp p; memset(&p, 0, sizeof p);
  for (int b = 0; b < 4; b++) { p.bc[b] = bc4[b]; p.qs[b] = qs4[b]; }
  p.mqSum = mqSum; p.qSum = qSum; p.deletions = deletions; p.insertions = insertions;
  p.insQual = insQual; p.delQual = delQual;
  const uint8_t* s = ins_bytes;
  for (int32_t k = 0; k < n_ins; k++) { olist_push(&p.ins, s, ins_len[k]); s += ins_len[k]; }
  s = del_bytes;
  for (int32_t k = 0; k < n_del; k++) { olist_push(&p.del, s, del_len[k]); s += del_len[k]; }
  BaseCall bc; baseCall(&p, cfg, &bc);
  out->base = (uint8_t)bc.base; out->alt = (uint8_t)bc.alt; out->q = (int32_t)bc.q;
  uint16_t f = 0;
  if (bc.homo) f |= PL_CF_HOMO;
  if (bc.highConf) f |= PL_CF_HIGHCONF;
  if (bc.called) f |= PL_CF_CALLED;
  if (bc.indel) f |= PL_CF_INDEL;
  if (bc.homoIndel) f |= PL_CF_HOMOINDEL;
  if (bc.isIns) f |= PL_CF_INSERTION;
  if (bc.isDel) f |= PL_CF_DELETION;
  if (bc.score > 0) f |= PL_CF_SCOREPOS;
  out->flags = f;
  *indel_len = bc.indelLen;
  *score = bc.score;
  free(p.ins.v); free(p.del.v);
}

/* ------------------------------------------------------ Assembler graph */
/* Assembler.addReads + buildGraph (Assembler.scala:46-85,131-153): every
 * K-mer of every read of length > K, on both strands (reverseComplement
 * maps A/C/G/T and keeps every other letter, Bases.scala:47-60; quals
 * reversed), piles the base after it with PileUp.add(base, qual, mq)
 * (PileUp.scala:75-84: A/C/G/T only, qual >= minQual; mq is the read's
 * MAPQ).  K-mers of depth >= minDepth get kGraph's edge (the called base)
 * and, when the call is not homozygous, altGraph's (the alternative base).
 * The pileups are aggregated by sorting the K-mer occurrences, which gives
 * the same sums as the HashMap (integer sums are order-free).  With 100,000
 * or more such reads Assembler.addRead prunes the map (:76), which this
 * restatement does not do: it returns PL_E_INVAL instead.
 * Output: the K-mers in byte order, K bytes each; next[i] the kGraph base,
 * alt[i] the altGraph base or 0. */
typedef struct { const uint8_t* k; uint8_t next; int8_t qual; int32_t mq; } KOcc;
static int32_t g_kmer_k;
static int cmp_kocc(const void* a, const void* b) {
  return memcmp(((const KOcc*)a)->k, ((const KOcc*)b)->k, (size_t)g_kmer_k);
}

int32_t orc_kmer_graph(const pl_config* cfg, const pl_read_soa* R, int32_t K, int32_t minDepth,
                       uint8_t** kmers, uint8_t** next, uint8_t** alt, int64_t* n_out) {
  static const char NIB[] = "=ACMGRSVTWYHKDBN";
  *kmers = NULL; *next = NULL; *alt = NULL; *n_out = 0;
  if (K < 1 || K > 48) return PL_E_INVAL;
  int64_t nocc = 0, nreads = 0, chars = 0;
  for (int64_t i = 0; i < R->n_reads; i++) {
    const int64_t L = R->seq_off[i + 1] - R->seq_off[i];
    if (L > K) { nreads++; nocc += 2 * (L - K); chars += 2 * L; }
  }
  if (nreads >= 100000) return PL_E_INVAL;
  uint8_t* strands = (uint8_t*)malloc((size_t)chars + 1);
  KOcc* occ = (KOcc*)malloc(sizeof(KOcc) * (size_t)(nocc + 1));
  int64_t co = 0, no = 0, ex = 0;
  int8_t* q = NULL; int64_t qcap = 0;
  for (int64_t i = 0; i < R->n_reads; i++) {
    const int64_t s0 = R->seq_off[i], L = R->seq_off[i + 1] - s0;
    while (ex < R->n_exotic && R->exo_read[ex] < i) ex++;
    const int exotic = ex < R->n_exotic && R->exo_read[ex] == i;
    if (L <= K) continue;
    uint8_t* fw = strands + co;
    uint8_t* rc = fw + L;
    co += 2 * L;
    if (L > qcap) { qcap = L; q = (int8_t*)realloc(q, (size_t)qcap); }
    for (int64_t j = 0; j < L; j++) {                       /* getReadString */
      uint8_t c;
      if (exotic) {
        const int64_t k = R->exo_off[ex] + j;
        c = (uint8_t)NIB[(R->nib[k >> 1] >> ((k & 1) ? 0 : 4)) & 15];
      } else {
        const int64_t k = s0 + j;
        c = (uint8_t)"ACGT"[(R->bases2[k >> 2] >> (2 * (k & 3))) & 3];
      }
      fw[j] = c;
      q[j] = (R->flag[i] & PL_RF_NO_QUAL) ? (int8_t)cfg->default_qual : R->quals[s0 + j];
    }
    for (int64_t j = 0; j < L; j++) {                       /* Bases.reverseComplement */
      const uint8_t c = fw[L - 1 - j];
      rc[j] = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : c == 'T' ? 'A' : c;
    }
    const int32_t mq = R->mapq[i];
    for (int64_t o = 0; o + K < L; o++) {                   /* addToPileups, offsets 0 .. L-K-1 */
      occ[no++] = (KOcc){fw + o, fw[o + K], q[o + K], mq};
      occ[no++] = (KOcc){rc + o, rc[o + K], q[L - 1 - (o + K)], mq};
    }
  }
  free(q);
  g_kmer_k = K;
  qsort(occ, (size_t)no, sizeof(KOcc), cmp_kocc);
  uint8_t* ko = (uint8_t*)malloc((size_t)K * (size_t)(no + 1));
  uint8_t* nx = (uint8_t*)malloc((size_t)no + 1);
  uint8_t* al = (uint8_t*)malloc((size_t)no + 1);
  int64_t n = 0;
  for (int64_t a = 0; a < no;) {
    int64_t b = a + 1;
    while (b < no && memcmp(occ[a].k, occ[b].k, (size_t)K) == 0) b++;
    PileUp p; memset(&p, 0, sizeof p);
    for (int64_t t = a; t < b; t++) {                       /* PileUp.add */
      const uint8_t c = occ[t].next;
      const int bi = c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : -1;
      if (bi >= 0 && occ[t].qual >= cfg->min_qual) {
        const int32_t mq1 = occ[t].mq + 1;
        p.bc[bi] += 1;
        p.qs[bi] += (int64_t)jmul(occ[t].qual, mq1);
        p.mqSum = jadd(p.mqSum, mq1);
        p.qSum = jadd(p.qSum, occ[t].qual);
      }
    }
    if (pu_depth(&p) >= minDepth) {                         /* buildGraph */
      BaseCall bc; baseCall(&p, cfg, &bc);
      memcpy(ko + (size_t)K * (size_t)n, occ[a].k, (size_t)K);
      nx[n] = (uint8_t)"ACGTN"[bc.base];
      al[n] = bc.homo ? 0 : (uint8_t)"ACGT"[bc.alt];
      n++;
    }
    a = b;
  }
  free(occ); free(strands);
  *kmers = ko; *next = nx; *alt = al; *n_out = n;
  return PL_OK;
}
void orc_free(void* p) { free(p); }
