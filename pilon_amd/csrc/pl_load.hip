// pl_load.hip — resident read sets built on the device from BAM record
// bytes (included at the end of pl_engine.hip).  The host inflates the BGZF
// blocks (pl_bam_stream_next of libpilonbam) straight into pinned staging
// buffers; each slice of whole records and its record starts go to the
// device as they come; at the end the records are decoded there, one
// thread per record, into the per-reference arrays a pl_read_soa holds
// (the same values pl_bam_open produces on the host, BamFile.scala:38-60,
// 108-148 through htsjdk's SAMRecord getters), and the read sets are
// prepared as pl_readset_upload_many prepares them.
//
// A reference's records are one run of the file: the BAM is sorted by
// reference in header order and by position within each (samtools sort),
// with the records of no reference last.

namespace {

__device__ __forceinline__ uint32_t ld_u32(const uint8_t* p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint32_t ld_u16(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }

// The CG:B,I tag of a record's tag area [p, e) (offsets from the record's
// start): word count and first word, or false.  Tag value sizes follow
// SAMv1 section 4.2.4.
__device__ bool find_cg_dev(const uint8_t* r, uint32_t p, uint32_t e, uint32_t* count, uint32_t* words) {
  while (p + 3 <= e) {
    const uint8_t t0 = r[p], t1 = r[p + 1], ty = r[p + 2];
    p += 3;
    uint32_t sz = 0;
    if (ty == 'A' || ty == 'c' || ty == 'C') sz = 1;
    else if (ty == 's' || ty == 'S') sz = 2;
    else if (ty == 'i' || ty == 'I' || ty == 'f') sz = 4;
    else if (ty == 'Z' || ty == 'H') {
      while (p < e && r[p]) p++;
      if (p >= e) return false;
      p++;
      continue;
    } else if (ty == 'B') {
      if (p + 5 > e) return false;
      const uint8_t sub = r[p];
      const uint32_t cnt = ld_u32(r + p + 1);
      uint32_t es = 0;
      if (sub == 'c' || sub == 'C') es = 1;
      else if (sub == 's' || sub == 'S') es = 2;
      else if (sub == 'i' || sub == 'I' || sub == 'f') es = 4;
      else return false;
      if (t0 == 'C' && t1 == 'G' && (sub == 'I' || sub == 'i')) {
        if ((uint64_t)p + 5 + (uint64_t)cnt * 4 > e) return false;
        *count = cnt;
        *words = p + 5;
        return true;
      }
      const uint64_t len = 5 + (uint64_t)cnt * es;
      if ((uint64_t)p + len > e) return false;
      sz = (uint32_t)len;
    } else {
      return false;
    }
    p += sz;
  }
  return false;
}

__device__ __forceinline__ bool acgt_nib(uint32_t v) { return v == 1 || v == 2 || v == 4 || v == 8; }
__device__ __forceinline__ uint32_t code_of_nib(uint32_t v) { return v == 1 ? 0u : v == 2 ? 1u : v == 4 ? 2u : 3u; }

struct LoadArgs {
  const uint8_t* R;        // record bytes of every slice, in file order
  const int64_t* ro;       // [N] record starts in R
  int64_t N;
  int64_t R_len;           // bytes in R
  int32_t nref;
  // per record
  int32_t* ref;
  int32_t* pos1;
  uint32_t* ncig;
  uint32_t* lseq;
  uint32_t* cigo;          // CIGAR words from the record's start (the body, or the CG tag's)
  uint32_t* exo;           // 1: a non-ACGT SEQ code
  uint32_t* nibn;          // nibble bytes of an exotic record
  // [0] lengths past block_size, [1] refID outside the dictionary, [2] references out of header order or
  // records of no reference before placed ones, [3] positions out of order, [4] a record not whole in the
  // bytes pushed; INT32_MAX: none
  int32_t* err;
  int32_t* first;          // [nref] first record of each reference, -1: none
  int32_t* nplaced;        // first record of no reference (N: none)
  // scans
  const uint64_t* cigscan;
  const uint64_t* seqscan;
  const uint32_t* exoscan;
  const uint64_t* nibscan;
  // per reference
  const int64_t* fr;       // [nref + 1] placed records of the references before it
  const int64_t* b2base;   // [nref] byte offset of its 2-bit bases (a multiple of 256)
  // outputs
  int32_t* pos;
  int32_t* tlen;
  uint16_t* flag;
  uint8_t* mapq;
  int64_t* cigar_off;      // reference r's n_r + 1 entries at fr[r] + r
  uint32_t* cigar;
  int64_t* seq_off;
  uint32_t* b2w;           // 2-bit bases as 32-bit words
  uint8_t* quals;
  int64_t* exo_read;
  int64_t* exo_off;
  uint8_t* nib;
};

// Each record's facts: lengths checked against block_size, the long CIGAR
// of a kSmN placeholder resolved through its CG tag, a non-ACGT code found.
__global__ void __launch_bounds__(256) k_ld_facts(LoadArgs A) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < A.N; k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = A.ro[k];
    // nothing is read before the record is known to lie whole in the bytes pushed
    bool whole = o >= 0 && o + 36 <= A.R_len;
    if (whole) {
      const uint32_t b = ld_u32(A.R + o);
      whole = b >= 32 && o + 4 + (int64_t)b <= A.R_len;
    }
    if (!whole) {
      atomicMin(&A.err[4], (int32_t)k);
      A.ref[k] = -1; A.pos1[k] = 0; A.ncig[k] = 0; A.lseq[k] = 0; A.cigo[k] = 0; A.exo[k] = 0; A.nibn[k] = 0;
      continue;
    }
    const uint8_t* r = A.R + o;
    const uint32_t bs = ld_u32(r), l_name = r[12], nc_raw = ld_u16(r + 16), ls = ld_u32(r + 20);
    const uint64_t body = 36 + (uint64_t)l_name + 4 * (uint64_t)nc_raw;
    if (ls > bs || body + ((uint64_t)ls + 1) / 2 + ls > 4 + (uint64_t)bs) {
      atomicMin(&A.err[0], (int32_t)k);
      A.ref[k] = -1; A.pos1[k] = 0; A.ncig[k] = 0; A.lseq[k] = 0; A.cigo[k] = 0; A.exo[k] = 0; A.nibn[k] = 0;
      continue;
    }
    const int32_t rf = (int32_t)ld_u32(r + 4);
    uint32_t nc = nc_raw, cig = 36 + l_name;
    if (nc == 2 && ld_u32(r + cig) == ((ls << 4) | 4u) && (ld_u32(r + cig + 4) & 15u) == 3u) {
      uint32_t cnt = 0, w = 0;
      if (find_cg_dev(r, (uint32_t)(body + (ls + 1) / 2 + ls), 4 + bs, &cnt, &w)) { nc = cnt; cig = w; }
    }
    const uint8_t* seq = r + body;
    bool ex = false;
    for (uint32_t j = 0; j < ls / 2; j++) {
      const uint32_t b = seq[j];
      ex = ex || !acgt_nib(b >> 4) || !acgt_nib(b & 15u);
    }
    if (ls & 1) ex = ex || !acgt_nib(seq[ls / 2] >> 4);
    A.ref[k] = rf;
    A.pos1[k] = (int32_t)ld_u32(r + 8) + 1;
    A.ncig[k] = nc;
    A.lseq[k] = ls;
    A.cigo[k] = cig;
    A.exo[k] = ex ? 1u : 0u;
    A.nibn[k] = ex ? (ls + 1) / 2 : 0u;
    if (rf < -1 || rf >= A.nref) atomicMin(&A.err[1], (int32_t)k);
  }
}

// Order checks and each reference's first record.
__global__ void __launch_bounds__(256) k_ld_runs(LoadArgs A) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < A.N; k += (int64_t)gridDim.x * blockDim.x) {
    const int32_t rf = A.ref[k];
    const int32_t eff = rf < 0 ? INT32_MAX : rf;
    const int32_t prev = k > 0 ? A.ref[k - 1] : INT32_MIN;
    if (k > 0) {
      const int32_t peff = prev < 0 ? INT32_MAX : prev;
      if (eff < peff) atomicMin(&A.err[2], (int32_t)k);
      else if (rf >= 0 && rf == prev && A.pos1[k] < A.pos1[k - 1]) atomicMin(&A.err[3], (int32_t)k);
    }
    if (rf >= 0 && rf < A.nref && (k == 0 || prev != rf)) A.first[rf] = (int32_t)k;
    if (rf < 0 && (k == 0 || prev >= 0)) atomicMin(A.nplaced, (int32_t)k);
  }
}

// Every placed record into its reference's arrays, one wave per record: the
// lanes copy its CIGAR words and QUAL bytes, build its 2-bit words (sixteen
// bases each; a word that holds bases of another read too is OR-ed
// atomically) or copy its nibbles, lane 0 writes the header fields.
__global__ void __launch_bounds__(256) k_ld_decode(LoadArgs A, int64_t n_placed) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t k = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < n_placed; k += nwaves) {
    const int32_t rf = A.ref[k];
    const int64_t f = A.fr[rf], i = k - f;
    const uint8_t* r = A.R + A.ro[k];
    const uint32_t l_name = r[12], nc_raw = ld_u16(r + 16), ls = A.lseq[k], nc = A.ncig[k];
    const uint8_t* seq = r + 36 + l_name + 4 * nc_raw;
    const uint8_t* qual = seq + (ls + 1) / 2;
    const bool ex = A.exo[k] != 0;
    const uint64_t c0 = A.cigscan[f], s0 = A.seqscan[f];
    if (lane == 0) {
      const bool noqual = ls == 0 || qual[0] == 0xFF;
      const int32_t next_ref = (int32_t)ld_u32(r + 24);
      A.pos[k] = A.pos1[k];
      A.tlen[k] = (int32_t)ld_u32(r + 32);
      A.mapq[k] = r[13];
      A.flag[k] = (uint16_t)((ld_u16(r + 18) & 0x0FFFu) | (noqual ? PL_RF_NO_QUAL : 0u) |
                             (next_ref == rf ? PL_RF_MATE_SAME_REF : 0u) | (ex ? PL_RF_EXOTIC : 0u));
      A.cigar_off[f + rf + i + 1] = (int64_t)(A.cigscan[k + 1] - c0);
      A.seq_off[f + rf + i + 1] = (int64_t)(A.seqscan[k + 1] - s0);
    }
    const uint8_t* cw = r + A.cigo[k];
    uint32_t* co = A.cigar + A.cigscan[k];
    for (uint32_t j = lane; j < nc; j += 64) co[j] = ld_u32(cw + 4 * j);
    uint8_t* qo = A.quals + A.seqscan[k];
    for (uint32_t j = lane; j < ls; j += 64) qo[j] = qual[j];
    if (!ex) {
      const int64_t sb = (int64_t)(A.seqscan[k] - s0), se = sb + ls;
      if (se > sb) {
        uint32_t* w0 = A.b2w + A.b2base[rf] / 4;
        const int64_t wf = sb >> 4, wl = (se - 1) >> 4;
        for (int64_t wi = wf + lane; wi <= wl; wi += 64) {
          const int64_t p0 = wi * 16 > sb ? wi * 16 : sb, p1 = wi * 16 + 16 < se ? wi * 16 + 16 : se;
          uint32_t w = 0;
          for (int64_t p = p0; p < p1; p++) {
            const uint32_t j = (uint32_t)(p - sb);
            const uint32_t b = seq[j >> 1];
            w |= code_of_nib((j & 1) ? (b & 15u) : (b >> 4)) << (2 * (uint32_t)(p & 15));
          }
          if (p0 == wi * 16 && p1 == wi * 16 + 16) w0[wi] = w;
          else atomicOr(&w0[wi], w);
        }
      }
    } else {
      const uint32_t e = A.exoscan[k];
      if (lane == 0) {
        A.exo_read[e] = i;
        A.exo_off[e] = 2 * (int64_t)(A.nibscan[k] - A.nibscan[f]);
      }
      uint8_t* no = A.nib + A.nibscan[k];
      const uint32_t nbytes = (ls + 1) / 2;
      for (uint32_t j = lane; j < nbytes; j += 64) {
        uint8_t v = seq[j];
        if ((ls & 1) && j == nbytes - 1) v &= 0xF0;
        no[j] = v;
      }
    }
  }
}

// Add a slice's byte offset to its record starts.
__global__ void __launch_bounds__(256) k_ld_rebase(int64_t* ro, int64_t n, int64_t base) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    ro[k] += base;
}

// The scan values at each reference's first record and past its last.
__global__ void __launch_bounds__(256) k_ld_table(LoadArgs A, uint64_t* tab) {
  const int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r > A.nref) return;
  const int64_t f = A.fr[r];
  tab[8 * (int64_t)r + 0] = A.cigscan[f];
  tab[8 * (int64_t)r + 1] = A.seqscan[f];
  tab[8 * (int64_t)r + 2] = A.exoscan[f];
  tab[8 * (int64_t)r + 3] = A.nibscan[f];
}

inline unsigned ld_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 1 << 16)); }

// 64-bit exclusive prefix sums of a u32 array (out[n] = the total)
pl_status scan_u32_to_u64(Lane* L, const uint32_t* in, int64_t n, uint64_t* out, VcfBufs& B) {
  const int64_t tiles = std::max<int64_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE);
  uint64_t* tsum;
  if (pl_status s = B.get(&tsum, 8 * (size_t)tiles)) return s;
  k_scan64_tiles<<<dim3((unsigned)tiles), dim3(256), 0, L->stream>>>(in, n, tsum);
  k_scan64_top<<<dim3(1), dim3(1024), 0, L->stream>>>(tsum, tiles);
  k_scan64_apply<<<dim3((unsigned)tiles), dim3(256), 0, L->stream>>>(in, n, tsum, out);
  HIPCHK(hipGetLastError());
  return PL_OK;
}

}  // namespace

struct pl_loader {
  pl_ctx* c = nullptr;
  int64_t cap = 0, rcap = 0;             // staging bytes and record starts per slice
  uint8_t* sb[2] = {nullptr, nullptr};
  int64_t* so[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int cur = 0;
  uint8_t* bytes = nullptr;              // device record bytes
  int64_t bytes_cap = 0, bytes_len = 0;
  int64_t* offs = nullptr;               // device record starts
  int64_t offs_cap = 0, n_rec = 0;
  ~pl_loader() {
    for (int k = 0; k < 2; k++) {
      if (ev[k]) { (void)hipEventSynchronize(ev[k]); (void)hipEventDestroy(ev[k]); }
      if (sb[k]) (void)hipHostFree(sb[k]);
      if (so[k]) (void)hipHostFree(so[k]);
    }
    if (bytes) (void)hipFree(bytes);
    if (offs) (void)hipFree(offs);
  }
};

namespace {

pl_status loader_finish(pl_loader* ld, int32_t nref, const int32_t* contig_ids, int32_t bam_type, int32_t* ids) {
  pl_ctx* c = ld->c;
  Lane* L = &c->main;
  hipStream_t st = L->stream;
  const int64_t N = ld->n_rec;
  HIPCHK(hipStreamSynchronize(st));
  VcfBufs B;
  LoadArgs A;
  std::memset(&A, 0, sizeof A);
  A.R = ld->bytes;
  A.ro = ld->offs;
  A.N = N;
  A.nref = nref;
  const size_t n1 = (size_t)N + 1;
  pl_status s;
  if ((s = B.get(&A.ref, 4 * n1)) || (s = B.get(&A.pos1, 4 * n1)) || (s = B.get(&A.ncig, 4 * n1)) ||
      (s = B.get(&A.lseq, 4 * n1)) || (s = B.get(&A.cigo, 4 * n1)) || (s = B.get(&A.exo, 4 * n1)) ||
      (s = B.get(&A.nibn, 4 * n1)) || (s = B.get(&A.err, 20)) || (s = B.get(&A.first, 4 * (size_t)(nref + 1))) ||
      (s = B.get(&A.nplaced, 4)))
    return s;
  // the scans' inputs hold a zero past the last record
  HIPCHK(hipMemsetAsync(A.ncig + N, 0, 4, st));
  HIPCHK(hipMemsetAsync(A.lseq + N, 0, 4, st));
  HIPCHK(hipMemsetAsync(A.exo + N, 0, 4, st));
  HIPCHK(hipMemsetAsync(A.nibn + N, 0, 4, st));
  const int32_t none[5] = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
  const int32_t nn = (int32_t)N;
  A.R_len = ld->bytes_len;
  HIPCHK(hipMemcpyAsync(A.err, none, 20, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(A.nplaced, &nn, 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(A.first, 0xFF, 4 * (size_t)(nref + 1), st));
  if (N > 0) {
    k_ld_facts<<<dim3(ld_grid(N)), dim3(256), 0, st>>>(A);
    HIPCHK(hipGetLastError());
    k_ld_runs<<<dim3(ld_grid(N)), dim3(256), 0, st>>>(A);
    HIPCHK(hipGetLastError());
  }
  uint64_t *cigscan, *seqscan, *nibscan;
  uint32_t* exoscan;
  if ((s = B.get(&cigscan, 8 * n1)) || (s = B.get(&seqscan, 8 * n1)) || (s = B.get(&nibscan, 8 * n1)) ||
      (s = B.get(&exoscan, 4 * n1)))
    return s;
  if ((s = scan_u32_to_u64(L, A.ncig, N, cigscan, B)) || (s = scan_u32_to_u64(L, A.lseq, N, seqscan, B)) ||
      (s = scan_u32_to_u64(L, A.nibn, N, nibscan, B)))
    return s;
  HIPCHK(L->tiles.ensure(sizeof(uint32_t) * (size_t)std::max<int64_t>(1, (N + SCAN_TILE - 1) / SCAN_TILE)));
  if ((s = scan_u32(L, A.exo, exoscan, N))) return s;
  A.cigscan = cigscan; A.seqscan = seqscan; A.nibscan = nibscan; A.exoscan = exoscan;
  int32_t err[5], nplaced = 0;
  std::vector<int32_t> first((size_t)nref + 1);
  HIPCHK(hipMemcpyAsync(err, A.err, 20, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(&nplaced, A.nplaced, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(first.data(), A.first, 4 * (size_t)(nref + 1), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (err[4] != INT32_MAX) return fail(PL_E_INVAL, "BAM record %d: not whole in the slices pushed", err[4]);
  if (err[0] != INT32_MAX) return fail(PL_E_INVAL, "BAM record %d: name, CIGAR, SEQ and QUAL exceed block_size", err[0]);
  if (err[1] != INT32_MAX) return fail(PL_E_INVAL, "BAM record %d: refID outside the dictionary", err[1]);
  if (err[2] != INT32_MAX)
    return fail(PL_E_INVAL, "BAM record %d: references out of header order, or a record of no reference before "
                "placed ones (the device loader reads a BAM sorted as samtools sorts)", err[2]);
  if (err[3] != INT32_MAX) return fail(PL_E_INVAL, "BAM record %d: BAM is not coordinate-sorted", err[3]);
  // each reference's run: placed records before it, its count
  std::vector<int64_t> fr((size_t)nref + 1, 0);
  {
    int64_t at = 0;
    for (int32_t r = 0; r < nref; r++) {
      fr[(size_t)r] = at;
      if (first[(size_t)r] >= 0) {
        if (first[(size_t)r] != at) return fail(PL_E_INVAL, "reference %d's records are not one run", r);
        int64_t end = nplaced;
        for (int32_t q = r + 1; q < nref; q++)
          if (first[(size_t)q] >= 0) { end = first[(size_t)q]; break; }
        at = end;
      }
    }
    fr[(size_t)nref] = at;
    if (at != nplaced) return fail(PL_E_INVAL, "placed records outside the references' runs");
  }
  int64_t* dfr;
  if ((s = B.get(&dfr, 8 * fr.size()))) return s;
  HIPCHK(hipMemcpyAsync(dfr, fr.data(), 8 * fr.size(), hipMemcpyHostToDevice, st));
  A.fr = dfr;
  uint64_t* dtab;
  if ((s = B.get(&dtab, 64 * (size_t)(nref + 1)))) return s;
  k_ld_table<<<dim3((unsigned)((nref + 1 + 255) / 256)), dim3(256), 0, st>>>(A, dtab);
  HIPCHK(hipGetLastError());
  std::vector<uint64_t> tab(8 * (size_t)(nref + 1));
  HIPCHK(hipMemcpyAsync(tab.data(), dtab, 64 * (size_t)(nref + 1), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // the arena of every reference's arrays
  auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
  auto T = [&](int32_t r, int k) { return tab[8 * (size_t)r + (size_t)k]; };
  const int64_t P = nplaced;
  const uint64_t tot_cig = T(nref, 0), tot_seq = T(nref, 1), tot_exo = T(nref, 2), tot_nib = T(nref, 3);
  std::vector<int64_t> b2base((size_t)nref + 1, 0);
  size_t b2 = 0;
  for (int32_t r = 0; r < nref; r++) {
    b2base[(size_t)r] = (int64_t)b2;
    b2 = align(b2 + (size_t)((T(r + 1, 1) - T(r, 1) + 3) / 4) + 64 + 16);
  }
  const bool lr = bam_type == PL_BAM_NANOPORE || bam_type == PL_BAM_PACBIO;
  size_t at = 0;
  auto take = [&](size_t bytes) { const size_t a = at; at = align(at + bytes + 64 + 16); return a; };
  const size_t o_pos = take(4 * (size_t)P), o_tlen = take(4 * (size_t)P), o_flag = take(2 * (size_t)P),
               o_mapq = take((size_t)P), o_coff = take(8 * (size_t)(P + nref)), o_cig = take(4 * (size_t)tot_cig),
               o_soff = take(8 * (size_t)(P + nref)), o_b2 = take(b2), o_q = take((size_t)tot_seq),
               o_exr = take(8 * (size_t)tot_exo), o_exo = take(8 * (size_t)tot_exo), o_nib = take((size_t)tot_nib),
               o_lrt = lr ? take((size_t)P) : 0;
  HIPCHK(hipSetDevice(c->device));
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(at, 256)));
  std::shared_ptr<void> arena(p, [](void* q) { (void)hipFree(q); });
  uint8_t* U = (uint8_t*)p;
  HIPCHK(hipMemsetAsync(U, 0, std::max<size_t>(at, 256), st));
  if (lr) HIPCHK(hipMemsetAsync(U + o_lrt, bam_type == PL_BAM_NANOPORE ? PL_LR_NANOPORE : PL_LR_PACBIO, (size_t)P, st));
  int64_t* db2;
  if ((s = B.get(&db2, 8 * b2base.size()))) return s;
  HIPCHK(hipMemcpyAsync(db2, b2base.data(), 8 * b2base.size(), hipMemcpyHostToDevice, st));
  A.b2base = db2;
  A.pos = (int32_t*)(U + o_pos); A.tlen = (int32_t*)(U + o_tlen); A.flag = (uint16_t*)(U + o_flag);
  A.mapq = U + o_mapq; A.cigar_off = (int64_t*)(U + o_coff); A.cigar = (uint32_t*)(U + o_cig);
  A.seq_off = (int64_t*)(U + o_soff); A.b2w = (uint32_t*)(U + o_b2); A.quals = U + o_q;
  A.exo_read = (int64_t*)(U + o_exr); A.exo_off = (int64_t*)(U + o_exo); A.nib = U + o_nib;
  if (P > 0) {
    k_ld_decode<<<dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>((P + 3) / 4, 1 << 16))), dim3(256), 0, st>>>(A, P);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipStreamSynchronize(st));
  // the read sets of the references kept, prepared as uploaded ones are
  std::vector<std::unique_ptr<ReadSet>> rs;
  std::vector<int64_t> ncig, nb;
  std::vector<int32_t> which;
  for (int32_t r = 0; r < nref; r++) {
    ids[r] = -1;
    if (contig_ids[r] < 0) continue;
    const int64_t f = fr[(size_t)r], n = fr[(size_t)r + 1] - f;
    std::unique_ptr<ReadSet> R(new ReadSet());
    R->contig = contig_ids[r];
    R->n = n;
    R->has_frags = bam_type == PL_BAM_FRAGS;
    R->has_nano = bam_type == PL_BAM_NANOPORE;
    R->shared[0] = arena;
    RSDev& d = R->d;
    std::memset(&d, 0, sizeof d);
    d.n = n;
    d.pos = (const int32_t*)(U + o_pos) + f;
    d.tlen = (const int32_t*)(U + o_tlen) + f;
    d.flag = (const uint16_t*)(U + o_flag) + f;
    d.mapq = (const uint8_t*)(U + o_mapq) + f;
    d.cigar_off = (const int64_t*)(U + o_coff) + f + r;
    d.cigar = (const uint32_t*)(U + o_cig) + T(r, 0);
    d.seq_off = (const int64_t*)(U + o_soff) + f + r;
    d.bases2 = U + o_b2 + b2base[(size_t)r];
    d.quals = (const int8_t*)(U + o_q) + T(r, 1);
    d.exo_read = (const int64_t*)(U + o_exr) + T(r, 2);
    d.exo_off = (const int64_t*)(U + o_exo) + T(r, 2);
    d.nib = U + o_nib + T(r, 3);
    d.n_exotic = (int64_t)(T(r + 1, 2) - T(r, 2));
    if (lr) d.lrt = U + o_lrt + f;
    rs.push_back(std::move(R));
    ncig.push_back((int64_t)(T(r + 1, 0) - T(r, 0)));
    nb.push_back((int64_t)(T(r + 1, 1) - T(r, 1)));
    which.push_back(r);
  }
  std::vector<int32_t> got(rs.size());
  if ((s = prep_sets(c, rs, ncig, nb, got.data()))) return s;
  for (size_t k = 0; k < which.size(); k++) ids[which[k]] = got[k];
  return PL_OK;
}

}  // namespace

extern "C" {

pl_status pl_loader_open(pl_ctx* c, int64_t record_bytes, int64_t slice_bytes, pl_loader** out) {
  if (!c || !out || record_bytes < 0 || slice_bytes < 65536) return fail(PL_E_INVAL, "bad loader arguments");
  if (c->serving) return fail(PL_E_BUSY, "a server is open on this context: its read sets are fixed");
  *out = nullptr;
  std::unique_ptr<pl_loader> ld(new pl_loader());
  ld->c = c;
  ld->cap = slice_bytes;
  ld->rcap = slice_bytes / 36 + 1;
  HIPCHK(hipSetDevice(c->device));
  for (int k = 0; k < 2; k++) {
    if (hipHostMalloc((void**)&ld->sb[k], (size_t)ld->cap, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ld->so[k], 8 * (size_t)ld->rcap, hipHostMallocDefault) != hipSuccess)
      return fail(PL_E_OOM, "pinned staging of %lld bytes", (long long)slice_bytes);
    HIPCHK(hipEventCreateWithFlags(&ld->ev[k], hipEventDisableTiming));
  }
  ld->bytes_cap = record_bytes + 64;
  ld->offs_cap = record_bytes / 36 + 1;
  HIPCHK(hipMalloc((void**)&ld->bytes, (size_t)ld->bytes_cap + 64));
  HIPCHK(hipMalloc((void**)&ld->offs, 8 * (size_t)ld->offs_cap));
  *out = ld.release();
  return PL_OK;
}

pl_status pl_loader_staging(pl_loader* ld, uint8_t** bytes, int64_t* cap, int64_t** rec_off, int64_t* rec_cap) {
  if (!ld || !bytes || !cap || !rec_off || !rec_cap) return fail(PL_E_INVAL, "null argument");
  HIPCHK(hipEventSynchronize(ld->ev[ld->cur]));        // its last copy is done
  *bytes = ld->sb[ld->cur];
  *cap = ld->cap;
  *rec_off = ld->so[ld->cur];
  *rec_cap = ld->rcap;
  return PL_OK;
}

pl_status pl_loader_push(pl_loader* ld, int64_t len, int64_t n_rec) {
  if (!ld || len < 0 || n_rec < 0 || len > ld->cap || n_rec > ld->rcap) return fail(PL_E_INVAL, "bad slice");
  if (ld->bytes_len + len > ld->bytes_cap || ld->n_rec + n_rec > ld->offs_cap)
    return fail(PL_E_INVAL, "slices beyond the record bytes the loader was opened for");
  if (ld->n_rec + n_rec >= INT32_MAX) return fail(PL_E_INVAL, "more than 2^31 records");
  hipStream_t st = ld->c->main.stream;
  // the starts become offsets into all bytes pushed on the device (where
  // pl_loader_finish checks that every record lies whole in them)
  HIPCHK(hipMemcpyAsync(ld->bytes + ld->bytes_len, ld->sb[ld->cur], (size_t)len, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(ld->offs + ld->n_rec, ld->so[ld->cur], 8 * (size_t)n_rec, hipMemcpyHostToDevice, st));
  HIPCHK(hipEventRecord(ld->ev[ld->cur], st));
  if (n_rec > 0) {
    k_ld_rebase<<<dim3(ld_grid(n_rec)), dim3(256), 0, st>>>(ld->offs + ld->n_rec, n_rec, ld->bytes_len);
    HIPCHK(hipGetLastError());
  }
  ld->bytes_len += len;
  ld->n_rec += n_rec;
  ld->cur ^= 1;
  return PL_OK;
}

pl_status pl_loader_finish(pl_loader* ld, int32_t n_refs, const int32_t* contig_ids, int32_t bam_type, int32_t* ids) {
  if (!ld || n_refs < 0 || (n_refs > 0 && (!contig_ids || !ids))) return fail(PL_E_INVAL, "null argument");
  if (bam_type != PL_BAM_FRAGS && bam_type != PL_BAM_UNPAIRED && bam_type != PL_BAM_NANOPORE && bam_type != PL_BAM_PACBIO)
    return fail(PL_E_INVAL, "bam type %d of a resident read set is not PL_BAM_FRAGS, UNPAIRED, NANOPORE or PACBIO", bam_type);
  if ((bam_type == PL_BAM_NANOPORE || bam_type == PL_BAM_PACBIO) && !ld->c->pc.longread)
    return fail(PL_E_INVAL, "a long-read BAM needs config.longread (Pilon.longread)");
  for (int32_t r = 0; r < n_refs; r++)
    if (contig_ids[r] >= 0 && ld->c->contigs.find(contig_ids[r]) == ld->c->contigs.end())
      return fail(PL_E_INVAL, "contig %d (reference %d) not uploaded", contig_ids[r], r);
  try {
    return loader_finish(ld, n_refs, contig_ids, bam_type, ids);
  } catch (const std::bad_alloc&) {
    return fail(PL_E_OOM, "host allocation failed");
  }
}

void pl_loader_close(pl_loader* ld) { delete ld; }

}  // extern "C"
