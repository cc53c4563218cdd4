/*
 * pilon_bam.h — BAM reader that produces the engine's read batches
 * (libpilonbam.so, host C++, libdeflate).
 *
 * Replaces the reading half of Pilon's BamFile (pilon/BamFile.scala:38-60,
 * 108-148): htsjdk's SamReader over a coordinate-sorted BAM and the
 * SAMRecord getters the pileup path reads —
 *
 *   getAlignmentStart        pos + 1                   (PileUpRegion.scala:103)
 *   getInferredInsertSize    tlen                      (PileUpRegion.scala:65-72)
 *   getFlags                 flag                      (BamFile.scala:101-105, PileUpRegion.scala:107)
 *   getMappingQuality        mapq                      (PileUpRegion.scala:107,139-141)
 *   getCigar                 BAM CIGAR words, or the CG:B,I tag for long CIGARs
 *   getReadBases             SEQ nibbles "=ACMGRSVTWYHKDBN" (PileUpRegion.scala:112)
 *   getBaseQualities         QUAL bytes; 0xFF first byte = absent (PileUpRegion.scala:114-115)
 *   getReferenceIndex == getMateReferenceIndex        (PileUpRegion.scala:107)
 *
 * Every record placed on a reference (refID >= 0, mapped or not) is handed
 * over, in file order, as one pl_read_soa per reference.  validateRead and
 * the query window [start-10000, stop+10000] of BamFile.process are applied
 * per region by the engine (pl_region_*, pl_batch_run), so a reference's
 * reads serve every region of it.  Records with refID -1 are left out, as
 * queryOverlapping never returns them.
 *
 * BGZF blocks are inflated on n_threads host threads (CRC32 and ISIZE are
 * checked), records are delimited in one pass and decoded into the batches
 * in parallel.  The file is read in slices of about 64 MiB of BAM data (the
 * environment variable PILON_BAM_SLICE_KIB sets another size), so host
 * memory holds one slice plus the decoded batches.
 */
#ifndef PILON_BAM_H
#define PILON_BAM_H

#include <stdint.h>
#include "pilon_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pl_bam pl_bam;

typedef struct pl_bam_stats {
  int64_t compressed_bytes;   /* file size                                  */
  int64_t bam_bytes;          /* inflated BAM stream (header + records)     */
  int64_t blocks;             /* BGZF blocks, the end-of-file marker included */
  int64_t records;            /* alignment records                          */
  int64_t placed;             /* records with refID >= 0 (handed over)      */
  int64_t exotic;             /* placed records with a non-ACGT SEQ code    */
  double  inflate_s;          /* wall time in BGZF inflate                  */
  double  decode_s;           /* wall time delimiting and decoding records  */
  int32_t threads;
  int32_t eof_marker;         /* 1 if the file ends with the BGZF EOF block */
} pl_bam_stats;

/* Reads and decodes a whole coordinate-sorted BAM.  PL_E_INVAL for a file
 * that is not BGZF/BAM, a corrupt block (CRC32 or ISIZE mismatch), a record
 * whose lengths disagree with its block_size, or records out of coordinate
 * order; PL_E_OOM when host memory runs out.  n_threads <= 0 uses the
 * host's cores (at most 64). */
pl_status pl_bam_open(const char* path, int32_t n_threads, pl_bam** out);

/* Opens a coordinate-sorted BAM through its BAI index (SAMv1 section 5.2):
 * only the header and the index are read here; records are read per
 * region by pl_bam_query.  index NULL looks for <path>.bai, then for path
 * with ".bam" replaced by ".bai" (htsjdk's names).  Like Pilon
 * (BamFile.scala:55-60, "must be indexed BAM") it refuses a BAM without an
 * index: PL_E_INVAL. */
pl_status pl_bam_open_indexed(const char* path, const char* index, int32_t n_threads, pl_bam** out);

/* htsjdk's queryOverlapping(name, start, stop) as BamFile.process calls it
 * (BamFile.scala:118-119): the placed records of reference `ref` whose
 * span [alignment start, alignment end] -- one base for an unmapped read
 * or an alignment without reference bases -- overlaps [start, stop]
 * (1-based, closed; start <= 0 reads from the first base), in file order,
 * as one read batch.  Only the BGZF blocks of the index chunks that can
 * hold such records (reg2bins bins, the linear index's lower bound) are
 * read and inflated.  The arrays belong to the pl_bam and stay valid until
 * the next query or pl_bam_close.  Needs pl_bam_open_indexed. */
pl_status pl_bam_query(pl_bam* bam, int32_t ref, int64_t start, int64_t stop, pl_read_soa* out);

/* The sequence dictionary (@SQ order = refID). */
int32_t     pl_bam_n_refs(const pl_bam* bam);
const char* pl_bam_ref_name(const pl_bam* bam, int32_t ref);
int64_t     pl_bam_ref_len(const pl_bam* bam, int32_t ref);

/* The records of reference `ref`, sorted by pos, as a read batch for
 * pl_readset_upload / pl_region_add_reads.  The arrays belong to the
 * pl_bam and stay valid until pl_bam_close.  The flag word is the SAM flag
 * (bits 0x001-0x800) with PL_RF_NO_QUAL, PL_RF_MATE_SAME_REF and
 * PL_RF_EXOTIC added; non-ACGT reads keep their nibbles at an even offset. */
pl_status pl_bam_ref_reads(const pl_bam* bam, int32_t ref, pl_read_soa* out);   /* pl_bam_open only */

/* A BAM read as a stream of record bytes, for a consumer that decodes the
 * records itself (the device loader, pl_loader_* in pilon_hip.h): the
 * header is parsed at open (pl_bam_n_refs, pl_bam_ref_name, pl_bam_ref_len
 * answer as after pl_bam_open); each pl_bam_stream_next inflates the next
 * BGZF blocks into dst (cap bytes, at least 64 KiB more than the longest
 * record), after the bytes of the record the previous call cut, and hands
 * over the whole records there: *len bytes holding *n_rec records whose
 * starts (their block_size fields) go to rec_off (rec_cap entries; cap / 36
 * + 1 always suffice).  *len == 0 at the end of the file.  Records are
 * delimited as pl_bam_open does; their fields are not checked here. */
pl_status pl_bam_stream_open(const char* path, int32_t n_threads, pl_bam** out);
pl_status pl_bam_stream_next(pl_bam* bam, uint8_t* dst, int64_t cap, int64_t* len, int64_t* rec_off,
                             int64_t rec_cap, int64_t* n_rec);
/* A stream's inflated size (header included): the ISIZE fields of all its
 * BGZF blocks, which bounds the record bytes it hands over; -1 if not a
 * stream. */
int64_t pl_bam_stream_bytes(const pl_bam* bam);

void pl_bam_get_stats(const pl_bam* bam, pl_bam_stats* out);
void pl_bam_close(pl_bam* bam);

/* Message of the last failing pl_bam_* call on this thread. */
const char* pl_bam_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
