// pl_common.h — structures shared by the host side and the gfx950 kernels of
// libpilonhip.so.  Plain PODs; device pointers are raw.
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>
#include "../../include/pilon_hip.h"

#define PL_WAVE 64

// per-read info word computed once per read set (k_readset_prep)
#define RI_ADJMQ_MASK   0x000000ffu  // adjMq (PileUpRegion.scala:139-141)
#define RI_VALID        0x00000100u  // mq>=minMq && (!paired || proper && same ref) (:107)
#define RI_KEEP         0x00000200u  // passes BamFile.validateRead (BamFile.scala:101-105)
#define RI_OUTLIER      0x00000400u  // span beyond the window scan margins
#define RI_EDGE_CLIP    0x00000800u  // S op on a read without SEQ (Pilon throws)
#define RI_EXOTIC       0x00001000u
#define RI_NOQUAL       0x00002000u
#define RI_EVENTS       0x00004000u  // has an I/D op passing valid && trusted
#define RI_BAD          0x00008000u  // CIGAR/SEQ disagree, or op code > 8
#define RI_UNK_SHIFT    16           // number of 'P' ops, saturated at 255
#define RI_UNK_MASK     0x00ff0000u
#define RI_SIMPLE       0x01000000u  // one M/=/X op covering all of SEQ, 2-bit bases
#define RI_LR_SHIFT     25           // longReadType of the read's BAM (BamFile.scala:43-47): 0, 1 nanopore, 2 pacbio
#define RI_LR_MASK      0x06000000u
#define RI_SEG_SHIFT    27           // complex reads: number of segments (RMeta.op0 = first)
#define RI_SEG_MASK     0xf8000000u
#define RI_SEG_MAX      31           // more segments, or a long read: the read is walked instead
#define PL_LR_NANOPORE  1
#define PL_LR_PACBIO    2

// Margins beyond which a read is an outlier and is handled from the per-set
// outlier list instead of the sorted scan.
#define PL_MAX_LEAD 4096
#define PL_MAX_BACK 8192
// a read with this much sequence or more is walked as an outlier, and a read
// whose sequence starts this far past its tile's first read takes the
// complex path: together they keep tile-relative sequence offsets in int32
#define PL_MAX_SEQ (1 << 20)
#define PL_TILE_SPAN (1 << 30)

// Per-read facts derived once per read set (k_prep), 32 bytes.
struct __attribute__((aligned(16))) RMeta {
  int32_t pos;               // getAlignmentStart
  uint32_t info;             // RI_* bits and adjMq
  int32_t fp_lo, fp_hi;      // positions the read can touch (M, clips, slid indels); empty if lo > hi
  int32_t fr_lo, fr_hi;      // physCovIncr fragment [lo, hi); none if lo >= hi
  int32_t aend;              // getAlignmentEnd (0 for unmapped)
  uint32_t op0;              // first CIGAR word (0 if none)
};

// The part of physCovIncr a window needs from a read whose pileup footprint
// ends before the window and whose fragment starts before it: the fragment's
// end and size.  hi = INT32_MIN when the read adds no fragment (no fragment,
// not kept, malformed, or handled from the outlier list).
struct FragTail {
  int32_t hi;                // fragment end (exclusive)
  int32_t size;              // end - start, Int arithmetic
};

struct RSDev {               // one resident read set (one contig, sorted by pos)
  int64_t n;
  const int32_t* pos;
  const int32_t* tlen;
  const uint16_t* flag;
  const uint8_t* mapq;
  const int64_t* cigar_off;
  const uint32_t* cigar;
  const int64_t* seq_off;
  const uint8_t* bases2;
  const int8_t* quals;
  int64_t n_exotic;
  const int64_t* exo_read;
  const int64_t* exo_off;
  const uint8_t* nib;
  const uint8_t* lrt;        // [n] longReadType per read (null: all 0)
  RMeta* meta;               // [n] derived per-read facts
  const struct Seg* segs;    // segment table of the complex reads
  const uint8_t* ref;        // contig bytes
  int64_t ref_len;
  int64_t n_outlier;
  int64_t* outliers;         // ascending read indices
  int32_t* scal;             // [0]=max_back [1]=max_lead [2]=max_qspan [3]=n_outlier(int)
                             // [6]=max_back of the pileup footprints alone
  FragTail* ftail;           // [n] per read
};

struct JobDev {              // one region of a batch
  int32_t rs;                // read set index
  int32_t start, stop;       // 1-based inclusive
  int32_t qs, qe;            // BamFile query interval
  int32_t size;
  int32_t win0, nwin;        // windows of this job
  int64_t out;               // global index of position `start`
  int64_t rlo, rhi;          // reads that can overlap the query (count kernel)
};

struct WinDev {
  int32_t job;
  int32_t p0, p1;            // region-relative [p0, p1)
  int32_t order;             // position in the work queue
  int64_t lo, hi;            // reads whose pos is in [locus(p0)-back, locus(p1)+lead)
  int64_t plo;               // reads in [lo, plo) reach the window only through their fragments
  int64_t pad_;
};

// What the window kernel needs about one window and its region, in the
// order the work queue hands windows out: one 80-byte load per window
// instead of a chain of dependent loads (order -> window -> region).
struct __attribute__((aligned(16))) WinHead {
  int32_t wi, job, p0, p1;   // window index, region, region-relative [p0, p1)
  int64_t lo, hi, plo;       // read ranges (WinDev)
  int64_t out;               // the region's first batch position
  int32_t rs, start, stop, qs, qe;   // read set and JobDev bounds
  int32_t pad[3];
};

struct JobStat {             // device-written per job
  int64_t read_count;
  int64_t base_count;
  int64_t aligned;
  int64_t unknown_ops;
  int64_t err_read_edge;     // lowest read on which Pilon throws (INT64_MAX: none)
  int64_t err_read_inval;    // lowest malformed read (INT64_MAX: none)
  int32_t overflow;          // a column beyond PL_DEPTH_LIMIT
  int32_t min_depth;
  int32_t n_confirmed;
  int32_t cand_lo, cand_hi;  // range in the sorted candidate list
  int32_t pad;
  uint32_t hist[256];        // confirmed candidates by depth (255 = 255+)
};

// one piece of insertion/deletion evidence (PileUp.addInsertion/addDeletion)
struct IndelRec {
  uint32_t key;              // (gpos << 1) | kind, kind 0 = insertion, 1 = deletion
  int32_t len;
  int16_t qual;              // quals(readOffset) before normalisation
  uint16_t mq1;              // indelMq + 1
  uint32_t hash;             // FNV-1a of the (rotated) string
  int32_t rot;               // insertion: right rotations applied
  uint32_t flags;            // bit0: source is nib stream; bit1: contains 'N'
  int64_t src;               // insertion: first base (2-bit or nibble index); deletion: contig offset
};

// a change candidate (SNP/AMB from the window kernel, INS/DEL from the indel pass)
struct Cand {
  uint32_t gpos;
  int32_t depth;             // count + deletions before any Pass-1 mutation
  int32_t kind;              // pl_disposition
  int32_t len;               // indel length
  int64_t rec;               // winning IndelRec index (INS/DEL)
};

struct KConf {               // the knobs kernels read (Pilon.scala:52-73)
  int32_t flank, min_mq, min_qual, default_qual, min_min_depth, old_indel, fixamb;
  int32_t nonpf, duplicates;
  double min_depth;
};

struct EngineArgs {
  KConf cf;
  const RSDev* rs;
  const JobDev* jobs;
  int32_t n_jobs;
  const WinDev* wins;
  int32_t n_wins;
  int32_t W;                 // window size (positions)
  int32_t* queue;            // persistent work queue head
  const int32_t* worder;     // window processing order
  const WinHead* heads;      // [n_wins] window heads in processing order
  int32_t* pile;             // [PL_NFIELDS][G]
  pl_call* calls;            // [G]
  int64_t G;
  JobStat* st;
  IndelRec* rec; uint32_t* rec_n; uint32_t rec_cap;
  Cand* cand; uint32_t* cand_n; uint32_t cand_cap;
  uint32_t* defer; uint32_t* defer_n; uint32_t defer_cap;
  uint32_t* rkey; uint32_t* rval;   // sorted record keys/indices
  uint32_t* mutated; uint32_t* mutated_n; uint32_t mutated_cap;
  int32_t* overflow;                // capacity overflow flag
  struct WideCol* wide; uint32_t* wide_n; uint32_t wide_cap;   // exact sums of wide columns
  int32_t* win_wide;                // [2 * n_wins]: each window's first wide entry and count
  uint32_t* deep;                   // set when some region's minDepth is above the confirmed histogram
};
