/*
 * pilon_hip.h — C ABI of the MI355X pileup engine (libpilonhip.so).
 *
 * This is the drop-in boundary for Pilon's pileup-construction path.  Every
 * entry point below replaces one piece of the Scala surface of Pilon 1.24:
 *
 *   PileUpRegion(name, start, stop)        pilon/PileUpRegion.scala:26
 *   PileUpRegion.addRead(r, refBases, lr)  pilon/PileUpRegion.scala:102-220
 *   BamFile.process / validateRead         pilon/BamFile.scala:101-148
 *   PileUpRegion.postProcess               pilon/PileUpRegion.scala:226-229
 *   GenomeRegion.postProcess (Pass 1)      pilon/GenomeRegion.scala:214-272
 *   PileUp fields + PileUp.baseCall        pilon/PileUp.scala:25-257
 *   GenomeRegion.identifyAndFixIssues      pilon/GenomeRegion.scala:307-380,413
 *   GenomeRegion.fixIssues/writeChanges    pilon/GenomeRegion.scala:557-657
 *   GenomeRegion.writeVcf / Vcf            pilon/GenomeRegion.scala:623-643, pilon/Vcf.scala:28-163
 *   Tracks.standardTracks, makeTrack,      pilon/Tracks.scala:33-216,
 *   makeBedTrack; Pass 2, computeGc        pilon/GenomeRegion.scala:158-210,273-283,696-763
 *   Assembler.addReads / buildGraph        pilon/Assembler.scala:46-85,131-153
 *
 * Conventions: plain C types only; 0 = success; the library owns device
 * memory, the caller owns host memory; one pl_ctx per host thread; objects
 * are not thread-safe.  pl_last_error() returns a thread-local message for
 * the last failing call.
 */
#ifndef PILON_HIP_H
#define PILON_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum pl_status {
  PL_OK = 0,
  PL_E_INVAL = 1,     /* bad argument or malformed read (e.g. CIGAR/SEQ length mismatch) */
  PL_E_OOM = 2,       /* device or host allocation failed */
  PL_E_DEVICE = 3,    /* HIP runtime error or no usable GPU */
  PL_E_EDGE = 4,      /* input on which Pilon 1.24 itself throws (see below) */
  PL_E_OVERFLOW = 5,  /* a column of 2^23 or more bases: a packed device counter could carry */
  PL_E_BUSY = 6,      /* serving queue full */
  PL_E_REJECTED = 7   /* serving request larger than the device-memory budget */
} pl_status;

/* PL_E_EDGE mirrors the inputs on which the reference raises an exception
 * instead of producing output: an insertion or deletion left-normalised past
 * region.start (negative pileup index, PileUpRegion.scala:161,182) and a
 * soft clip on a read without SEQ (bases(rOff) out of range,
 * PileUpRegion.scala:201). */

/* ---------------------------------------------------------------- config */
/* The knobs the path reads from Pilon's global object (Pilon.scala:52-73). */
typedef struct pl_config {
  int32_t flank;          /* Pilon.flank = 10                  */
  int32_t min_mq;         /* Pilon.minMq = 0                   */
  int32_t min_qual;       /* Pilon.minQual = 0                 */
  int32_t default_qual;   /* Pilon.defaultQual = 10 (a Byte)   */
  double  min_depth;      /* Pilon.minDepth = 0.1              */
  int32_t min_min_depth;  /* Pilon.minMinDepth = 5             */
  int32_t old_indel;      /* Pilon.oldIndel = false            */
  int32_t fix_snps;       /* "snps" in --fix                   */
  int32_t fix_indels;     /* "indels" in --fix                 */
  int32_t fix_amb;        /* "amb" in --fix                    */
  int32_t iupac;          /* --iupac                           */
  int32_t longread;       /* Pilon.longread (any --nanopore/--pacbio BAM) */
  int32_t nonpf;          /* --nonpf                           */
  int32_t duplicates;     /* --duplicates                      */
} pl_config;

/* Defaults identical to Pilon 1.24 with "--fix bases". */
void pl_config_default(pl_config* cfg);

/* ------------------------------------------------------------ read batch */
/* Extra bits the host packer sets in the 16-bit flag word (above the SAM
 * flags, which stop at 0x800). */
#define PL_RF_NO_QUAL        0x2000u  /* BAM QUAL absent: use default_qual (PileUpRegion.scala:114-115) */
#define PL_RF_MATE_SAME_REF  0x4000u  /* refID == next refID (PileUpRegion.scala:107) */
#define PL_RF_EXOTIC         0x8000u  /* SEQ holds a non-ACGT code: bases come from nib */

/* Packed structure-of-arrays of reads of ONE contig, sorted by pos.  It is
 * what a JVM/C++ host produces from htsjdk SAMRecords (or BAM records):
 *   pos      = getAlignmentStart            tlen = getInferredInsertSize
 *   flag     = getFlags | PL_RF_*           mapq = getMappingQuality
 *   cigar    = BAM cigar words (len<<4|op)  bases/quals = getReadBases/getBaseQualities
 * Bases are 2-bit codes A=0 C=1 G=2 T=3; a read with any other base sets
 * PL_RF_EXOTIC and lists its 4-bit BAM codes ("=ACMGRSVTWYHKDBN") in nib. */
typedef struct pl_read_soa {
  int64_t         n_reads;
  const int32_t*  pos;        /* [n] 1-based alignment start, non-decreasing          */
  const int32_t*  tlen;       /* [n]                                                   */
  const uint16_t* flag;       /* [n]                                                   */
  const uint8_t*  mapq;       /* [n]                                                   */
  const int64_t*  cigar_off;  /* [n+1] prefix offsets into cigar                       */
  const uint32_t* cigar;      /* [cigar_off[n]]                                        */
  const int64_t*  seq_off;    /* [n+1] prefix offsets into the per-base streams        */
  const uint8_t*  bases2;     /* base k at byte k>>2, bits 2*(k&3)                     */
  const int8_t*   quals;      /* [seq_off[n]] Phred bytes as Java reads them (signed)  */
  int64_t         n_exotic;
  const int64_t*  exo_read;   /* [n_exotic] ascending read indices with PL_RF_EXOTIC   */
  const int64_t*  exo_off;    /* [n_exotic] first nibble of that read in nib           */
  const uint8_t*  nib;        /* nibble k at byte k>>1, high nibble when k is even     */
} pl_read_soa;

/* ------------------------------------------------------------- outputs */
/* PileUp counters, field-major: out[f*G + i] for position i of the region.
 * Order follows the field list of PileUp (PileUp.scala:26-39). */
enum pl_field {
  PL_F_COUNT_A = 0, PL_F_COUNT_C, PL_F_COUNT_G, PL_F_COUNT_T,   /* baseCount.sums */
  PL_F_QUAL_A, PL_F_QUAL_C, PL_F_QUAL_G, PL_F_QUAL_T,           /* qualSum.sums   */
  PL_F_MQSUM, PL_F_QSUM, PL_F_PHYSCOV, PL_F_INSERTSIZE, PL_F_BADPAIR,
  PL_F_DELETIONS, PL_F_DELQUAL, PL_F_INSERTIONS, PL_F_INSQUAL, PL_F_CLIPS,
  PL_NFIELDS
};

/* One PileUp.BaseCall plus the Pass-1 disposition (PileUp.scala:132-255,
 * GenomeRegion.scala:237-272), 8 bytes. */
typedef struct pl_call {
  uint8_t  base;   /* 0..3 = A,C,G,T; 4 = N (no bases)     BaseCall.base    */
  uint8_t  alt;    /* 0..3                                  BaseCall.altBase */
  uint16_t flags;  /* PL_CF_* below                                          */
  int32_t  q;      /* BaseCall.q = score / n                                 */
} pl_call;

#define PL_CF_HOMO        0x0001u  /* BaseCall.homo              */
#define PL_CF_HIGHCONF    0x0002u  /* q >= 10                     */
#define PL_CF_CALLED      0x0004u  /* base != 'N' || indel        */
#define PL_CF_INDEL       0x0008u  /* indel call present          */
#define PL_CF_HOMOINDEL   0x0010u  /* BaseCall.homoIndel          */
#define PL_CF_INSERTION   0x0020u  /* isInsertion                 */
#define PL_CF_DELETION    0x0040u  /* isDeletion                  */
#define PL_CF_SCOREPOS    0x0080u  /* score > 0                   */
#define PL_CF_DISP_SHIFT  8        /* 3-bit disposition           */
#define PL_CF_DISP_MASK   0x0700u
#define PL_CF_DELETED     0x0800u  /* GenomeRegion.deleted(i)     */
#define PL_CF_WIDE        0x1000u  /* a qualSum (a Long, BaseSum.scala:24) leaves int32: the
                                      pileup holds its low 32 bits, pl_last_wide_columns the sums */

enum pl_disposition { PL_D_NONE = 0, PL_D_CONFIRMED = 1, PL_D_SNP = 2,
                      PL_D_INS = 3, PL_D_DEL = 4, PL_D_AMB = 5 };

typedef struct pl_region_stats {
  int64_t read_count;   /* PileUpRegion.readCount                         */
  int64_t base_count;   /* PileUpRegion.baseCount                         */
  int64_t coverage;     /* roundDiv(baseCount, size)                      */
  int32_t min_depth;    /* GenomeRegion.minDepth                          */
  int32_t n_confirmed;  /* confirmed count ("Confirmed N of M bases")     */
  int32_t snps, amb, ins, ins_bases, dels, del_bases;
  int32_t n_unknown_ops;/* 'P' ops that print "unknown cigar op"          */
  int64_t aligned_bases;/* M/=/X bases with locus in the region (bench unit) */
} pl_region_stats;

/* ------------------------------------------------------------- context */
typedef struct pl_ctx pl_ctx;
typedef struct pl_region pl_region;

pl_status pl_ctx_create(const pl_config* cfg, int device, pl_ctx** out);
void      pl_ctx_destroy(pl_ctx* ctx);
const char* pl_last_error(void);

/* Upload a whole contig (raw FASTA bytes, case kept) and keep it resident.
 * Replaces the refBases argument of addRead (PileUpRegion.scala:102) and
 * GenomeRegion.contigBases (GenomeRegion.scala:35). */
pl_status pl_ref_upload(pl_ctx* ctx, int32_t contig_id, const char* name,
                        const uint8_t* ascii, int64_t len);

/* Upload a sorted batch of one contig's reads and keep it resident; returns
 * an id usable by pl_region_use_readset and by the batch/serving forms. */
pl_status pl_readset_upload(pl_ctx* ctx, int32_t contig_id,
                            const pl_read_soa* reads, int32_t* readset_id);
/* The same for a set that is one BAM of a given PL_BAM_* type (below):
 * frags, unpaired, nanopore or pacbio (long-read rules as for
 * pl_region_add_reads_typed; a region of a nanopore or pacbio set has no
 * frags BAM, so excludeMotifs applies).  A jumps BAM is refused here
 * (PL_E_INVAL): it only matters beside other BAMs of a region, through
 * pl_region_add_reads_typed.  pl_readset_upload is PL_BAM_FRAGS. */
pl_status pl_readset_upload_typed(pl_ctx* ctx, int32_t contig_id, const pl_read_soa* reads,
                                  int32_t bam_type, int32_t* readset_id);

/* Many read sets at once, as pl_readset_upload_typed would make each (one
 * per contig, e.g. every reference of a BAM): reads[i] of contig
 * contig_ids[i] with BAM type bam_types[i] (null: all PL_BAM_FRAGS); ids[i]
 * gets its id.  The host arrays go to the device through pinned staging
 * buffers filled by host threads, and the sets share device allocations, so
 * the cost is that of the bytes rather than of the sets.  No counterpart in
 * the reference: the resident form's loader (BamFile.process reads a BAM's
 * records per region, BamFile.scala:108-148). */
/* Waits until the context's device has finished all work queued on it
 * (hipDeviceSynchronize on its device).  Every pl_* call that returns
 * results returns synchronised; this is for callers that time or bracket
 * work.  No counterpart in the reference. */
pl_status pl_ctx_synchronize(pl_ctx* ctx);

pl_status pl_readset_upload_many(pl_ctx* ctx, int32_t n_sets, const int32_t* contig_ids,
                                 const pl_read_soa* const* reads, const int32_t* bam_types, int32_t* ids);

/* Resident read sets built on the device from a BAM's record bytes: the
 * host inflates (libpilonbam's pl_bam_stream_next) into the loader's pinned
 * staging buffers, the slices go to the device as they come, and
 * pl_loader_finish decodes every record there into the per-reference
 * arrays pl_bam_open would produce on the host, then prepares the sets as
 * pl_readset_upload_many does.  record_bytes bounds the record bytes of all
 * slices (the stream's inflated size); slice_bytes is the staging size.
 * The BAM must be sorted as samtools sorts (references in header order,
 * records of no reference last: PL_E_INVAL otherwise).  contig_ids[r] is
 * the context contig of BAM reference r (-1: skipped); ids[r] its set id
 * (-1 when skipped).  No counterpart in the reference: the resident form's
 * loader (BamFile.process reads a BAM's records per region,
 * BamFile.scala:108-148).  Use:
 *   pl_loader_open; loop { pl_loader_staging; pl_bam_stream_next into it;
 *   pl_loader_push } until the stream ends; pl_loader_finish;
 *   pl_loader_close. */
typedef struct pl_loader pl_loader;
pl_status pl_loader_open(pl_ctx* ctx, int64_t record_bytes, int64_t slice_bytes, pl_loader** out);
pl_status pl_loader_staging(pl_loader* ld, uint8_t** bytes, int64_t* cap, int64_t** rec_off, int64_t* rec_cap);
pl_status pl_loader_push(pl_loader* ld, int64_t len, int64_t n_rec);
pl_status pl_loader_finish(pl_loader* ld, int32_t n_refs, const int32_t* contig_ids, int32_t bam_type, int32_t* ids);
void pl_loader_close(pl_loader* ld);

/* --------------------------------------------------- plain region form */
/* new PileUpRegion(name, start, stop): 1-based inclusive, stop <= len.   */
pl_status pl_region_open(pl_ctx* ctx, int32_t contig_id, int32_t start,
                         int32_t stop, pl_region** out);
/* BamFile.process for one BAM: the region keeps the reads that overlap
 * [max(start-10000,0), min(stop+10000,len)] and pass validateRead, and adds
 * each with addRead.  Callable repeatedly (one call per BAM or batch). */
pl_status pl_region_add_reads(pl_region* reg, const pl_read_soa* reads);
/* The same for one BAM of a given type (Pilon's --frags, --jumps,
 * --unpaired, --nanopore, --pacbio; pl_region_add_reads is PL_BAM_FRAGS).
 * The depth that "jumps" BAMs add is left out of fragment coverage, and so
 * of copy number, the delta/dip tracks and their Pilon.bed and VCF regions
 * (processBam, GenomeRegion.scala:287-300).  Nanopore and pacbio BAMs are
 * unpaired BAMs of longReadType 1 and 2 (BamFile.scala:43-47): their reads
 * pile with addRead's long-read rules (indelMq capped at 8, insertions and
 * deletions dropped at homopolymers, nanopore CC_GG bases at quality 0;
 * PileUpRegion.scala:120-190) and need config.longread.  When a region has
 * no frags BAM and config.longread is set, excludeMotifs keeps changes at
 * homopolymer starts (and, with a nanopore BAM, CC_GG middles) out of the
 * fixes and counts (GenomeRegion.scala:219,326,432-439).  Every call adds
 * its BAM's type to the region, an empty one included. */
#define PL_BAM_FRAGS    0
#define PL_BAM_JUMPS    1
#define PL_BAM_UNPAIRED 2
#define PL_BAM_NANOPORE 3
#define PL_BAM_PACBIO   4
pl_status pl_region_add_reads_typed(pl_region* reg, const pl_read_soa* reads, int32_t bam_type);
/* Same, for a resident read set. */
pl_status pl_region_use_readset(pl_region* reg, int32_t readset_id);
/* PileUpRegion.postProcess + GenomeRegion.postProcess Pass 1.  pileup
 * (PL_NFIELDS*G int32) and calls (G) are optional caller buffers. */
pl_status pl_region_finish(pl_region* reg, int32_t* pileup, pl_call* calls,
                           pl_region_stats* stats);
/* identifyAndFixIssues (snps/indels) + writeChanges + fixed bases.  The
 * library allocates *changes_text (the .changes lines) and *bases; free
 * them with pl_free.  new_name/offset are writeChanges' arguments (new_name
 * NULL: the contig's polished name, as pl_pilon_name). */
pl_status pl_region_fix(pl_region* reg, const char* new_name, int64_t offset,
                        char** changes_text, int64_t* changes_len,
                        uint8_t** bases, int64_t* bases_len);
void      pl_region_close(pl_region* reg);
void      pl_free(void* p);

/* ------------------------------------------------- batched region form */
/* One region of a batch: the GenomeFile.processRegions loop body
 * (GenomeFile.scala:101-120) for thousands of regions in one launch.     */
typedef struct pl_region_desc {
  int32_t contig_id;
  int32_t start, stop;     /* 1-based inclusive                              */
  int32_t readset_id;      /* resident reads of that contig                  */
} pl_region_desc;

typedef struct pl_batch_opts {
  int32_t window;          /* positions per workgroup window; 0 = autotuned  */
  int32_t block;           /* threads per workgroup; 0 = autotuned           */
  int32_t persistent;      /* 1 = persistent work-queue launch               */
  int32_t use_graph;       /* 1 = replay a captured hipGraph                 */
  int32_t copy_out;        /* 1 = copy pileups/calls to the host buffers     */
} pl_batch_opts;

/* Runs every region of the batch on the device.  pileup/calls may be NULL
 * (results stay resident); otherwise they hold sum(G) entries in region
 * order.  stats has n entries. */
pl_status pl_batch_run(pl_ctx* ctx, const pl_region_desc* regions, int32_t n,
                       const pl_batch_opts* opts, int32_t* pileup,
                       pl_call* calls, pl_region_stats* stats);

/* n_passes more passes of the context's last batch, which pl_batch_run ran
 * with use_graph: its captured pipeline replayed back to back with no host
 * round trip between passes, each pass's evidence counters copied to pinned
 * memory and checked once the passes (at most 64 at a time) have drained
 * (PL_E_OOM if one overflowed: run the batch with pl_batch_run again).  The
 * results of the last pass stay resident, as after pl_batch_run;
 * pl_last_timing's pipeline time is the mean per pass.  No counterpart in
 * the reference: for callers that run the same resident batch repeatedly
 * (benchmarks, re-polishing loops) without a host wait per pass. */
pl_status pl_batch_repeat(pl_ctx* ctx, int32_t n_passes);

/* pl_batch_run without allocating: no hipMalloc, hipFree or other
 * allocating routine runs in this call.  It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * Every device buffer of the batch pipeline is carved from a scratch array
 * of 192 MiB that the library holds on each GPU it runs on until the
 * process ends, so a batch may hold at most 2^20 positions (PL_E_INVAL
 * beyond; PL_E_OOM if its buffers still do not fit); window/block 0 mean
 * 512/256 (no autotuning) and use_graph must be 0.  Inputs are the
 * context's resident read sets, outputs the caller's host buffers, as for
 * pl_batch_run.  All calls on one GPU share its array, which keeps the
 * last call's contents: calls that overlap on one GPU (from any thread or
 * context) race.  pl_last_batch_info and the VCF/track/fix entry points
 * keep describing the last pl_batch_run, not this call. */
pl_status pl_batch_run_noalloc(pl_ctx* ctx, const pl_region_desc* regions, int32_t n, const pl_batch_opts* opts,
                               int32_t* pileup, pl_call* calls, pl_region_stats* stats);

/* --------------------------------------------- local reassembly (f4) */
/* GapFiller.fillGap / fixBreak (GapFiller.scala:39-262) given the reads the
 * caller recruited for the break (BamFile.recruitFlankReads and
 * recruitBadMates stay with the caller: they need read names and the
 * whole-BAM insert statistics).  The K-mer graph of the reads is built on
 * the device (Assembler.addReads + buildGraph); Assembler's path search
 * (Assembler.scala:163-280) and the joins run on the host. */
typedef struct pl_asm_opts {
  int32_t k;              /* Assembler.K = 47                                   */
  int32_t min_depth;      /* Assembler.minDepth = 5                              */
  int32_t break_radius;   /* GapFiller.breakRadius: max(3K, mean insert size of
                             the region's frags BAMs, rounded)                  */
  int32_t tr_safe;        /* Pilon.trSafe = 1                                    */
  int32_t multi_closure;  /* Pilon.multiClosure = 0                              */
  int32_t fix_breaks;     /* Pilon.fixBreaks = 0 (--fix breaks)                  */
  int32_t gap_margin;     /* Pilon.gapMargin = 100000                            */
  int32_t min_gap;        /* Pilon.minGap = 10                                   */
  int32_t fix_gaps;       /* "gaps" in --fix (default on): fill the gaps         */
  int32_t fix_local;      /* "local" in --fix (default on): fix possible breaks  */
} pl_asm_opts;
void pl_asm_opts_default(pl_asm_opts* opts);

/* The fix GapFiller returns: (start, ref, patch), (0, "", "") for none, and
 * the tandem repeat its path search noted.  The three strings are
 * malloc'ed, NUL-terminated; free each with pl_free. */
typedef struct pl_asm_fix {
  int32_t start;
  char* ref;
  int64_t ref_len;
  char* patch;
  int64_t patch_len;
  char* tandem_repeat;
} pl_asm_fix;

/* One gap (is_gap = 1, fillGap) or break (fixBreak) [break_start,
 * break_stop] of region [region_start, region_stop] of a contig.
 * region_bases: the region's current bases (GenomeRegion.bases, SNP fixes
 * applied), region_stop - region_start + 1 bytes; contig: the whole contig
 * as read (contigBases), contig_len bytes.  Returns PL_E_EDGE where Pilon
 * 1.24 throws (a substring that reaches outside the region). */
pl_status pl_assemble_break(pl_ctx* ctx, const pl_read_soa* reads, const char* region_bases, const char* contig,
                            int64_t contig_len, int32_t region_start, int32_t region_stop, int32_t break_start,
                            int32_t break_stop, int32_t is_gap, const pl_asm_opts* opts, pl_asm_fix* out);

/* identifyAndFixIssues' local reassembly for one finished region
 * (GenomeRegion.scala:382-410) as Pilon runs it with --nostrays: for each
 * gap (a run of 10 or more N, not within 100 of the region's ends) fillGap,
 * then for each possible break (GenomeRegion.possibleBreaks) fixBreak, each
 * over the reads recruited from the read sets given (BamFile.readsInRegion
 * over flankRegion: the valid reads overlapping the break widened by the
 * set's maxInsertSize, frags sets first, then the unpaired ones unless
 * config.longread) and the region's bases after its SNP fixes.  A fix kept
 * (a gap's with start > 0, a break's when its ref or patch exceeds 10
 * bases) joins bigFixList: pl_region_fix then applies it after the
 * snps/indels fixes and lists it in the changes, and pl_region_bed adds each
 * logged attempt's outcome line (Tracks.scala:211-213); pl_region_vcf
 * writes a Vcf.writeFixRecord record at each kept fix that is not a small
 * one, the loci it replaces embedded, and a pl_region_bed after it counts
 * those loci as changed (Pilon writes Pilon.bed after the VCF).  sets are the read
 * sets the region was built from, with their PL_BAM_* types and
 * maxInsertSize (BamFile.maxInsertSize); a jumps set is PL_E_INVAL
 * (recruitBadMates pairs reads by name).  opts->break_radius is
 * GapFiller.breakRadius for this region.  *n_fixes = the fixes kept.
 * Calling it again replaces the region's reassembly. */
pl_status pl_region_reassemble(pl_region* reg, const pl_read_soa* const* sets, const int32_t* set_types,
                               const int32_t* max_insert, int32_t n_sets, const pl_asm_opts* opts, int32_t* n_fixes);

/* The lines GenomeRegion.logFix prints for the last pl_region_reassemble
 * ("fix gap: ctg:801-900 801 -0 +100 ClosedGap", "# fix break: ... 0 -0 +0
 * NoSolution"), newline-terminated; free with pl_free. */
pl_status pl_region_reassembly_log(pl_region* reg, char** text, int64_t* len);

/* Device time of the last pl_batch_run's pipeline in milliseconds, measured
 * with HIP events on the library's stream, plus the share of the window
 * kernel alone. */
pl_status pl_last_timing(pl_ctx* ctx, double* pipeline_ms, double* window_ms);

/* ------------------------------------------------------------------ VCF */
/* GenomeRegion.writeVcf (GenomeRegion.scala:623-643) for every region of
 * the context's last batch (pl_batch_run, or the region that pl_region_finish
 * ran), in batch order: Vcf.writeRecord (Vcf.scala:74-163) per locus, a
 * deleted locus as an embedded "Del" record, an indel call followed by its
 * second record, and Vcf.writeDup (Vcf.scala:194-202) before the first
 * locus of each duplicationEvents region (copy number above 1 over more
 * than 10,000 loci, GenomeRegion.scala:731-737).  Formatted on the device
 * from the pileup, calls and indel evidence the batch left there; *text is
 * malloc'ed (pl_free), *len bytes plus a terminating NUL.  Reassembly
 * records (writeFixRecord) are not produced: those fixes are outside this
 * path.  device_ms (may be NULL) = device time of the formatting passes
 * and of the copy-number passes behind the duplication records. */
#define PL_VCF_DIPLOID 0x2u   /* --diploid: no "Amb" filter              */
#define PL_VCF_QE      0x4u   /* --vcfqe: QE= (qualSum) instead of QP=  */
pl_status pl_batch_vcf(pl_ctx* ctx, uint32_t flags, char** text, int64_t* len, double* device_ms);

/* The same text handed to the caller in order, a piece at a time, instead
 * of one allocation: sink(user, data, n) receives each piece (data is valid
 * during the call only) and returns 0 to go on, anything else to stop
 * (PL_E_INVAL).  Pieces come from pinned buffers the context keeps, so a
 * sink that writes them to a file or socket overlaps the next piece's DMA.
 * *len (may be NULL) = the bytes handed over. */
typedef int32_t (*pl_text_sink)(void* user, const char* data, int64_t n);
pl_status pl_batch_vcf_stream(pl_ctx* ctx, uint32_t flags, pl_text_sink sink, void* user, int64_t* len,
                              double* device_ms);

/* The same for one finished region (pl_region_finish).  A region fed by
 * pl_region_add_reads keeps its reads on the device until pl_region_close
 * for this; if the context has run other batches since, the region's batch
 * runs again first.  pl_batch_vcf after pl_region_finish of such a region
 * returns PL_E_INVAL. */
pl_status pl_region_vcf(pl_region* reg, uint32_t flags, char** text, int64_t* len, double* device_ms);

/* Vcf.writeHeader (Vcf.scala:28-70): the header lines for the contigs
 * (names, lengths), with the caller's date (yyyyMMdd), source (Pilon's
 * version string), command line and reference URI. */
pl_status pl_vcf_header(const char* const* names, const int64_t* lengths, int32_t n, uint32_t flags,
                        const char* date, const char* source, const char* command, const char* reference,
                        char** text, int64_t* len);

/* --------------------------------------------------------------- tracks */
/* The per-locus tracks of Tracks.standardTracks (Tracks.scala:33-56), in
 * that order, with Pass 2 of GenomeRegion.postProcess (copy number,
 * GenomeRegion.scala:273-283) and computeGc (GenomeRegion.scala:160-188)
 * for every region of the context's last batch.  The Pilon.bed issue track
 * is pl_batch_bed below. */
enum pl_track {
  PL_TRACK_CHANGES = 0,        /* Changes.wig            */
  PL_TRACK_UNCONFIRMED,        /* Unconfirmed.wig        */
  PL_TRACK_COPY_NUMBER,        /* CopyNumber.wig         */
  PL_TRACK_COVERAGE,           /* Coverage.wig           */
  PL_TRACK_BAD_COVERAGE,       /* BadCoverage.wig        */
  PL_TRACK_PCT_BAD,            /* PctBad.wig             */
  PL_TRACK_DELTA_COVERAGE,     /* DeltaCoverage.wig      */
  PL_TRACK_DIP_COVERAGE,       /* DipCoverage.wig        */
  PL_TRACK_PHYSICAL_COVERAGE,  /* PhysicalCoverage.wig   */
  PL_TRACK_CLIPPED_ALIGNMENTS, /* ClippedAlignments.wig  */
  PL_TRACK_WEIGHTED_QUAL,      /* WeightedQual.wig       */
  PL_TRACK_WEIGHTED_MQ,        /* WeightedMq.wig         */
  PL_TRACK_GC,                 /* GC.wig                 */
  PL_NTRACKS
};

/* values[t * G + i] = track t at position i of the batch (G = the batch's
 * positions, regions in batch order); n_values >= PL_NTRACKS * G.  values
 * may be NULL to run the passes without copying (device_ms still set). */
pl_status pl_batch_tracks(pl_ctx* ctx, int32_t* values, int64_t n_values, double* device_ms);
/* The same for one finished region (as pl_region_vcf). */
pl_status pl_region_tracks(pl_region* reg, int32_t* values, int64_t n_values, double* device_ms);

/* Tracks.makeTrack (Tracks.scala:160-177) over the last batch's regions:
 * "fixedStep chrom=<contig> start=<start> step=1" per region, then one value
 * per line, formatted on the device; PL_WIG_TRACK_LINE puts the file's
 * "track type=wiggle_0 ..." line first.  *text is malloc'ed (pl_free). */
#define PL_WIG_TRACK_LINE 0x1u   /* also: pl_batch_bed's track line */
pl_status pl_batch_wig(pl_ctx* ctx, int32_t track, uint32_t flags, char** text, int64_t* len, double* device_ms);
pl_status pl_region_wig(pl_region* reg, int32_t track, uint32_t flags, char** text, int64_t* len, double* device_ms);

/* Tracks.makeBedTrack (Tracks.scala:194-216) over the last batch's regions:
 * per region its unconfirmed ("?"), change ("X"), low-coverage ("LowCov"),
 * collapsed-repeat ("Copy#"), duplication, gap and break regions
 * (GenomeRegion.scala:696-763) as BED lines.  The per-locus tests run on the
 * device with the tracks; the runs are summarised on the host.  The
 * reassembly outcomes Pilon appends per region are not produced (no
 * reassembly on this path).  PL_WIG_TRACK_LINE puts the file's
 * "track description=..." line first.  *text is malloc'ed (pl_free). */
pl_status pl_batch_bed(pl_ctx* ctx, uint32_t flags, char** text, int64_t* len, double* device_ms);
pl_status pl_region_bed(pl_region* reg, uint32_t flags, char** text, int64_t* len, double* device_ms);

/* ----------------------------------------------------- assembler K-mers */
/* Assembler.addReads + buildGraph (Assembler.scala:46-85,131-153), the
 * K-mer pileups of Pilon's local reassembly: every K-mer of every read
 * longer than K, on both strands, piles the base after it (PileUp.add, with
 * the read's MAPQ); each K-mer of depth >= min_depth gets its kGraph edge
 * (the called base, next[i]) and, when the call is heterozygous, its
 * altGraph edge (alt[i], else 0).  *kmers holds *n K-mers of K bytes each,
 * in byte order.  K in [1, 48]; PL_E_INVAL for 100,000 or more reads longer
 * than K (Assembler.addRead prunes its map there, which this pass does not).
 * The three arrays are malloc'ed (pl_free). */
pl_status pl_kmer_graph(pl_ctx* ctx, const pl_read_soa* reads, int32_t K, int32_t min_depth, char** kmers,
                        uint8_t** next, uint8_t** alt, int64_t* n, double* device_ms);

/* The columns of the context's last batch (pl_batch_run, pl_region_finish)
 * whose qualSum of some base -- a Long in BaseSum (BaseSum.scala:24) --
 * leaves int32, in position order: their calls carry PL_CF_WIDE and the
 * pileup planes hold the low 32 bits of those sums.  pos is the position in
 * the batch (the region index for pl_region_finish).  *n = their count; out
 * (cap entries; NULL to count only) receives them.  A column needs about
 * 66,000 bases of quality 40 at MAPQ 60 for this. */
typedef struct pl_wide_col {
  int64_t pos;
  int64_t qual_sum[4];   /* BaseSum qualSum.sums for A, C, G, T */
} pl_wide_col;
pl_status pl_last_wide_columns(pl_ctx* ctx, pl_wide_col* out, int64_t cap, int64_t* n);

/* ------------------------------------------------- fixes, FASTA, regions */
/* GenomeFile.processRegions' output loop (GenomeFile.scala:135-162) over the
 * context's last batch: per region, writeChanges (GenomeRegion.scala:645-657)
 * after identifyAndFixIssues (snps/indels/amb, :307-380,413) and fixIssues
 * (:595-621); per element, writeFastaElement (GenomeFile.scala:79-82) of the
 * fixed bases of its regions concatenated.  An element is one entry of
 * GenomeFile.regions: a maximal run of consecutive regions of one contig
 * with the same element[i] (element NULL: the contig, as for the chunks of
 * contigRegions; one element per --targets entry otherwise).  Its newName is
 * the contig name + "_pilon", "|pilon" or "pilon" (a name holding '|', one
 * ending in '|'), and the .changes offsets of a region accumulate the length
 * changes of the element's earlier regions (:149-153).  *changes_text and
 * *fasta_text (either may be NULL) are malloc'ed (pl_free). */
pl_status pl_batch_fix(pl_ctx* ctx, const int32_t* element, int32_t n_element, char** changes_text,
                       int64_t* changes_len, char** fasta_text, int64_t* fasta_len);

/* GenomeFile.writeFastaElement (GenomeFile.scala:79-82): ">header", then the
 * sequence in lines of 80 (the last shorter), each ended by '\n'.  *text is
 * malloc'ed (pl_free). */
pl_status pl_fasta_element(const char* header, const uint8_t* seq, int64_t len, char** text, int64_t* text_len);

/* The polished contig name of GenomeFile.processRegions (GenomeFile.scala:
 * 137-140).  *out is malloc'ed (pl_free). */
pl_status pl_pilon_name(const char* name, char** out);

/* GenomeFile.contigRegions (GenomeFile.scala:67-74) with Pilon.chunkSize
 * (10,000,000 by default, Pilon.scala:52): nChunks = ceil(len / chunk),
 * regions [1 + k*size, min(len, (k+1)*size)] of size ceil(len / nChunks), in
 * Int arithmetic.  *n = the region count; starts/stops (cap entries; both
 * NULL to count only) receive them. */
pl_status pl_contig_regions(int64_t length, int32_t chunk_size, int32_t* starts, int32_t* stops, int32_t cap,
                            int32_t* n);

/* One --targets region: contig = index into the names given. */
typedef struct pl_target {
  int32_t contig;
  int32_t start, stop;   /* 1-based inclusive */
} pl_target;

/* GenomeFile.parseTargets (GenomeFile.scala:215-250): the lines of the file
 * named by targets when it reads and every line parses to regions, else the
 * string itself; each comma-separated "contig:start-stop" (or a bare contig
 * name: the whole contig) is one unchunked region, in order, duplicates
 * kept.  Where Pilon throws (an unknown contig, a coordinate beyond Int,
 * stop past the contig) or builds a region it cannot process (start < 1,
 * stop < start) this returns PL_E_INVAL.  *out is malloc'ed (pl_free). */
pl_status pl_parse_targets(const char* targets, const char* const* names, const int64_t* lengths, int32_t n_contigs,
                           pl_target** out, int32_t* n_out);

/* Sizes and HIP-event times of the last batch (pl_batch_run or
 * pl_region_finish). */
typedef struct pl_batch_info {
  int64_t positions;       /* G, positions piled                          */
  int32_t n_regions;
  int32_t n_windows;       /* workgroup windows of the window kernel      */
  int64_t n_records;       /* insertion/deletion evidence records         */
  int64_t n_candidates;    /* SNP/AMB/INS/DEL candidates before Pass 1    */
  int64_t n_deferred;      /* columns whose BaseCall needed indel strings */
  int32_t window;          /* window size used                            */
  int32_t persistent;
  int32_t block;           /* workgroup size used                         */
  int32_t use_graph;       /* 1 = the pipeline ran as a graph replay      */
  double  pipeline_ms;     /* whole device pipeline                       */
  double  window_ms;       /* window kernel alone; -1 after a graph replay */
} pl_batch_info;
pl_status pl_last_batch_info(pl_ctx* ctx, pl_batch_info* info);

/* ------------------------------------------------------- serving form */
/* A request loop over the contigs and read sets resident in the context.
 * Polish requests (one region each) arrive at any time.  Each lane is an
 * independent batch pipeline on its own stream, driven by a host thread; an
 * idle lane takes every waiting request that fits beside the requests in
 * flight (estimated device bytes against mem_budget, FIFO, up to max_batch)
 * and runs them as one batch, so requests that arrive while others are on
 * the GPU join the next free lane and batches overlap on the device
 * (continuous batching).  A request's result is identical alone or in any
 * batch.  Each request is GenomeFile.processRegions' loop body for one region
 * (GenomeFile.scala:101-120); the serving loop has no reference counterpart.
 * While a server is open the context's contigs and read sets are fixed
 * (uploads return PL_E_BUSY). */
typedef struct pl_serve_opts {
  int32_t lanes;           /* concurrent batches; 0 = 2 (at most 16)                  */
  int32_t max_batch;       /* requests per batch; 0 = 256                             */
  int64_t mem_budget;      /* device bytes for in-flight requests; 0 = 1/4 of free     */
  int32_t max_queue;       /* waiting requests before PL_E_BUSY; 0 = 65536            */
  int32_t window, block;   /* launch shape; 0 = 512 / 256                             */
  int32_t copy_out;        /* PL_SERVE_KEEP_* bits: what each result keeps for fetching */
} pl_serve_opts;

#define PL_SERVE_KEEP_PILEUP 0x1u  /* pileup and calls, for pl_serve_fetch               */
#define PL_SERVE_KEEP_FIX    0x2u  /* change list, for pl_serve_fetch_fix               */

typedef struct pl_serve_result {
  int64_t ticket;
  int64_t tag;             /* the caller's tag from pl_serve_submit                   */
  int32_t status;          /* pl_status of this request                               */
  int32_t batch_size;      /* requests in the batch it ran in                         */
  pl_region_stats stats;
  int64_t est_bytes;       /* the admission estimate                                  */
  double  queue_ms;        /* submit -> its batch started                             */
  double  latency_ms;      /* submit -> result ready                                  */
  double  batch_device_ms; /* device time of its batch (HIP events)                   */
} pl_serve_result;

typedef struct pl_server pl_server;
pl_status pl_serve_open(pl_ctx* ctx, const pl_serve_opts* opts, pl_server** out);
/* PL_E_REJECTED: the request alone exceeds mem_budget; PL_E_BUSY: queue full. */
pl_status pl_serve_submit(pl_server* srv, const pl_region_desc* req, int64_t tag, int64_t* ticket);
/* Up to max finished requests; waits up to wait_ms (-1 = until one finishes
 * or nothing is outstanding, 0 = no wait). */
pl_status pl_serve_poll(pl_server* srv, pl_serve_result* out, int32_t max, int32_t wait_ms, int32_t* n_out);
/* Pileup (PL_NFIELDS*G, field-major) and calls of a polled request of a
 * copy_out server; the kept copy is freed afterwards. */
pl_status pl_serve_fetch(pl_server* srv, int64_t ticket, int32_t* pileup, pl_call* calls);
/* The .changes lines and fixed bases of a polled request of a server kept
 * with PL_SERVE_KEEP_FIX: identifyAndFixIssues + fixIssues + writeChanges
 * (GenomeRegion.scala:307-380,595-657) of its region, new_name NULL = the
 * contig's polished name; the kept change list is freed afterwards.  Both
 * outputs are malloc'ed (pl_free). */
pl_status pl_serve_fetch_fix(pl_server* srv, int64_t ticket, const char* new_name, int64_t offset, char** changes_text,
                             int64_t* changes_len, uint8_t** bases, int64_t* bases_len);
/* Waits until every submitted request has finished. */
pl_status pl_serve_drain(pl_server* srv);
/* Drains the queue, stops the lanes and frees their device memory. */
void      pl_serve_close(pl_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* PILON_HIP_H */
