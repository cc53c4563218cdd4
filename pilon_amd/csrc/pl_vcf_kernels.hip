// pl_vcf_kernels.hip — the VCF records of a batch's regions, formatted on the
// device (GenomeRegion.writeVcf, GenomeRegion.scala:623-643, and
// Vcf.writeRecord, Vcf.scala:74-163).  Included at the end of
// pl_kernels.hip: it uses that file's wave helpers.
//
// Three passes over the windows of the last batch, all reading the final
// pileup, calls and sorted indel evidence the batch left in HBM:
//   k_vcf_mark    lists the columns whose call is an insertion or deletion
//                 (and not at region index 0, where writeRecord never
//                 prints the indel allele);
//   k_vcf_indel   finds each listed column's winning evidence record again
//                 (the strict majority of the indel pass);
//   k_vcf_lines   measures every locus's record lines; an exclusive
//                 64-bit scan of the lengths gives each locus its offset;
//   k_vcf_write   formats them again, staged in LDS, and stores them.

struct VcfArgs {
  EngineArgs A;
  const char* names;          // region names, concatenated
  const int64_t* name_off;    // [n_jobs + 1]
  uint32_t flags;             // PL_VCF_* (include/pilon_hip.h)
  uint32_t* ilist;            // indel-call columns
  uint32_t* ilist_n;
  uint32_t icap;
  int32_t* imap;              // column -> its index in ilist (listed columns only)
  int64_t* irec;              // [icap] winning record of each listed column
  uint32_t* len;              // [G] bytes of the column's lines
  const uint64_t* off;        // [G + 1] their offsets
  char* text;
  const uint8_t* embed;       // [G] or null: loci writeVcf marks deleted after a reassembly record
};

namespace {

// Text sink: counts bytes, and stores them when WRITE.
template <bool WRITE>
struct VOut {
  char* p;
  uint32_t n;
  __device__ void c(char x) {
    if (WRITE) p[n] = x;
    n++;
  }
  __device__ void s(const char* t) {
    for (; *t; t++) c(*t);
  }
  // a Long in decimal (Java's Long.toString)
  __device__ void l(int64_t v) {
    uint64_t u = v < 0 ? (uint64_t)(-(v + 1)) + 1u : (uint64_t)v;
    if (v < 0) c('-');
    if (u <= 0xffffffffull) {
      // the common case: 32-bit divisions by 10 (a multiply-high each)
      uint32_t w = (uint32_t)u;
      uint64_t packed = 0;
      int nd = 0;
      do {
        packed |= (uint64_t)(w % 10u) << (4 * nd);
        w /= 10u;
        nd++;
      } while (w);
      for (int k = nd - 1; k >= 0; k--) c((char)('0' + ((packed >> (4 * k)) & 15u)));
    } else if (u < 10000000000000000ull) {
      // up to 16 digits, four bits each, generated with constant divisions
      uint64_t packed = 0;
      int nd = 0;
      do {
        packed |= (u % 10u) << (4 * nd);
        u /= 10u;
        nd++;
      } while (u);
      for (int k = nd - 1; k >= 0; k--) c((char)('0' + ((packed >> (4 * k)) & 15u)));
    } else {
      uint64_t pw = 1;
      while (u / pw >= 10u) pw *= 10u;
      for (; pw; pw /= 10u) c((char)('0' + (u / pw) % 10u));
    }
  }
};

// Java's "%.2f" of a Float widened to Double (Vcf.scala:147): digits of the
// exact value rounded half up.  A widened float has at most 24 significant
// bits, so its shortest decimal form never sits on a tie its exact value
// misses, and the exact value decides the same digits.
template <bool WRITE>
__device__ void put_af(VOut<WRITE>& o, float af) {
  uint64_t q = 0;
  if (af > 0.0f) {
    const uint32_t bits = __float_as_uint(af);
    const int ex = (int)((bits >> 23) & 255u);              // af is normal: a ratio of ints >= 1/2^31
    const uint64_t m = (uint64_t)((bits & 0x7fffffu) | 0x800000u);
    const uint64_t t = m * 100u;                             // af * 100 = t * 2^(ex - 150)
    const int sh = 150 - ex;
    if (sh <= 0) q = t << -sh;
    else if (sh < 63) {
      q = t >> sh;
      if ((t & ((1ull << sh) - 1u)) >= (1ull << (sh - 1))) q++;
    }
  }
  o.l((int64_t)(q / 100u));
  o.c('.');
  o.c((char)('0' + (q / 10u) % 10u));
  o.c((char)('0' + q % 10u));
}

// character q of a winning indel record's string: an insertion's comes from
// the read (rec_char), a deletion's is the reference at its locus, raw bytes
// (PileUpRegion.scala:182: addDeletion(refBases.slice(dloc - 1, dloc + len - 1)))
__device__ char indel_char(const RSDev& R, const IndelRec& e, int32_t q) {
  return (e.key & 1u) ? (char)R.ref[e.src + q] : (char)rec_char(R, e, q);
}

struct VcfCol {
  Col c;
  int32_t physCov, badPair, clips;
};

__device__ VcfCol load_vcf_col(const int32_t* pile, int64_t G, int64_t g) {
  VcfCol v;
  v.c = load_col(pile, G, g);
  v.physCov = pile[(int64_t)PL_F_PHYSCOV * G + g];
  v.badPair = pile[(int64_t)PL_F_BADPAIR * G + g];
  v.clips = pile[(int64_t)PL_F_CLIPS * G + g];
  return v;
}

// One Vcf.writeRecord call (Vcf.scala:74-163) for region index i at global
// column g.  The indel string (callString) is rec's, when indelOk.
template <bool WRITE>
__device__ __forceinline__ void vcf_record(VOut<WRITE>& o, const VcfArgs& V, const JobDev& J, const RSDev& R, const char* name,
                           int64_t name_len, int32_t minDepth, const VcfCol& v, uint16_t cflags, int64_t rec,
                           int32_t i, bool embedded, bool indelOkArg) {
  const Col& c = v.c;
  const BaseFacts bf = base_facts(c);
  int o0, o1;
  order4(c.qSum > 0 ? c.qs : c.bc, o0, o1);
  const bool indelOk = indelOkArg && i > 0;
  const bool isIns = (cflags & PL_CF_INSERTION) != 0, isDel = (cflags & PL_CF_DELETION) != 0;
  const bool homoIndel = (cflags & PL_CF_HOMOINDEL) != 0;
  const char base = bf.base == 4 ? 'N' : (char)acgt_char(bf.base);
  const char altB = (char)acgt_char(bf.alt);
  const bool useStr = indelOk && (isIns || isDel) && rec >= 0;
  const IndelRec* e = useStr ? &V.A.rec[rec] : nullptr;
  const int32_t slen = useStr ? e->len : 1;
  const int32_t baseDP = (int32_t)pick4(c.qs, o0), altBaseDP = (int32_t)pick4(c.qs, o1);
  const int64_t depth = c.depth();
  int32_t loc = J.start + i;
  // kind: 0 deletion, 1 insertion, 2 homozygous base, 3 heterozygous base
  int kind;
  int32_t refDP, altDP;
  const char* callType;
  if (indelOk && !embedded && isDel) {
    kind = 0;
    loc -= 1;
    callType = homoIndel ? "1/1" : "0/1";
    const int32_t p = del_pct(c);
    refDP = 100 - p; altDP = p;
  } else if (indelOk && !embedded && isIns) {
    kind = 1;
    loc -= 1;
    callType = homoIndel ? "1/1" : "0/1";
    const int32_t p = ins_pct(c);
    refDP = 100 - p; altDP = p;
  } else if (bf.homo) {
    kind = 2;
    const char rb = (char)to_upper(R.ref[loc - 1]);
    if (rb == base || (!useStr && base == 'N')) { callType = "0/0"; refDP = baseDP; altDP = altBaseDP; }
    else { callType = "1/1"; refDP = altBaseDP; altDP = baseDP; }
  } else {
    kind = 3;
    const char rb = (char)to_upper(R.ref[loc - 1]);
    callType = "0/1";
    if (rb == base) { refDP = baseDP; altDP = altBaseDP; }
    else { refDP = altBaseDP; altDP = baseDP; }
  }
  const char rBase = (char)to_upper(R.ref[loc - 1]);
  // the called allele (cBase) and whether it prints as "."
  char cb1 = base;
  if (kind == 3 && rBase == base) cb1 = altB;
  // Vcf.scala:131: cBase is the anchor base for a deletion (ALT "." when it
  // is N), the anchor plus the insertion for an insertion (never "."), the
  // called base otherwise
  const bool dotAlt = kind >= 2 ? (cb1 == 'N' || cb1 == rBase) : (kind == 0 && rBase == 'N');
  const int ac = callType[0] == '0' ? (callType[2] == '0' ? 0 : 1) : 2;
  const int32_t dsum = j_add(refDP, altDP);
  const float af = (dsum > 0 && !dotAlt) ? (float)altDP / (float)dsum : 0.0f;
  const int64_t count = c.count();
  const int64_t qsum = c.qs[0] + c.qs[1] + c.qs[2] + c.qs[3];
  for (int64_t k = 0; k < name_len; k++) o.c(name[k]);
  o.c('\t'); o.l(loc); o.s("\t.\t");
  // REF
  o.c(rBase);
  if (kind == 0) for (int32_t q = 0; q < slen; q++) o.c(indel_char(R, *e, q));
  o.c('\t');
  // ALT
  if (kind == 0) o.c(dotAlt ? '.' : rBase);
  else if (kind == 1) { o.c(rBase); for (int32_t q = 0; q < slen; q++) o.c(indel_char(R, *e, q)); }
  else if (dotAlt) o.c('.');
  else o.c(cb1);
  o.c('\t');
  if (indelOk && isDel) o.c('.'); else o.l(bf.score);
  o.c('\t');
  // filters: prepended in writeRecord, so the last added prints first
  bool any = false;
  if (embedded) { o.s("Del"); any = true; }
  if (!(V.flags & PL_VCF_DIPLOID) && callType[0] == '0' && callType[2] == '1') { if (any) o.c(';'); o.s("Amb"); any = true; }
  if (depth < minDepth) { if (any) o.c(';'); o.s("LowCov"); any = true; }
  if (!any) o.s("PASS");
  o.s("\tDP="); o.l(embedded ? count : depth);
  o.s(";TD="); o.l(depth + v.badPair);
  o.s(";BQ="); o.l(round_div_l(qsum, round_div_l((int64_t)c.mqSum * count, depth)));   // PileUp.scala:64-67
  o.s(";MQ="); o.l(round_div_l((int64_t)c.mqSum - depth, depth));                     // PileUp.scala:70-72
  o.s(";QD="); o.l(bf.q);
  o.s(";BC=");
  for (int b = 0; b < 4; b++) { if (b) o.c(','); o.l(c.bc[b]); }
  if (V.flags & PL_VCF_QE) {                                                           // BaseSum.toString
    o.s(";QE=");
    for (int b = 0; b < 4; b++) { if (b) o.c(','); o.l(c.qs[b]); }
  } else {                                                                             // BaseSum.toStringPct
    o.s(";QP=");
    for (int b = 0; b < 4; b++) { if (b) o.c(','); o.l(qsum == 0 ? 0 : (100 * c.qs[b] + qsum / 2) / qsum); }
  }
  o.s(";PC="); o.l(v.physCov);
  o.s(";IC="); o.l(c.insertions);
  o.s(";DC="); o.l(c.deletions);
  o.s(";XC="); o.l(v.clips);
  o.s(";AC="); o.l(ac);
  o.s(";AF="); put_af(o, af);
  o.s("\tGT\t"); o.s(callType); o.c('\n');
}

// all records of one locus: writeRecord(i, deleted(i)) and, for an indel
// call, its second record (Vcf.scala:161)
template <bool WRITE>
__device__ __forceinline__ uint32_t vcf_locus(char* out, const VcfArgs& V, const JobDev& J, const RSDev& R, const char* name,
                              int64_t name_len, int32_t minDepth, int64_t g, int32_t i) {
  VOut<WRITE> o;
  o.p = out;
  o.n = 0;
  VcfCol v = load_vcf_col(V.A.pile, V.A.G, g);
  const uint16_t f = V.A.calls[g].flags;
  if (f & PL_CF_WIDE) wide_patch(v.c, V.A.wide, V.A.wide_cap, V.A.win_wide, J, V.A.W, g);
  const bool deleted = (f & PL_CF_DELETED) != 0 || (V.embed && V.embed[g]);
  int64_t rec = -1;
  if (i > 0 && (f & (PL_CF_INSERTION | PL_CF_DELETION))) rec = V.irec[V.imap[g]];
  vcf_record<WRITE>(o, V, J, R, name, name_len, minDepth, v, f, rec, i, deleted, true);
  if (i > 0 && (f & PL_CF_INDEL) && !deleted)
    vcf_record<WRITE>(o, V, J, R, name, name_len, minDepth, v, f, rec, i,
                      (f & PL_CF_DELETION) && (f & PL_CF_HOMOINDEL), false);
  return o.n;
}

}  // namespace

__global__ void __launch_bounds__(256) k_vcf_mark(VcfArgs V) {
  const WinDev win = V.A.wins[blockIdx.x];
  const JobDev& J = V.A.jobs[win.job];
  for (int p = win.p0 + threadIdx.x; p - threadIdx.x < win.p1; p += blockDim.x) {
    const bool act = p < win.p1;
    const int64_t g = J.out + p;
    const bool want = act && p > 0 && (V.A.calls[g].flags & (PL_CF_INSERTION | PL_CF_DELETION));
    const uint32_t k = wave_append(V.ilist_n, want);
    if (want && k < V.icap) { V.ilist[k] = (uint32_t)g; V.imap[g] = (int32_t)k; }   // the host re-runs when k >= icap
  }
}

// one wave per listed column: the winning record, as the indel pass found it
__global__ void __launch_bounds__(256) k_vcf_indel(VcfArgs V) {
  const EngineArgs& A = V.A;
  const uint32_t n = *V.ilist_n < V.icap ? *V.ilist_n : V.icap;
  const uint32_t nrec = *A.rec_n < A.rec_cap ? *A.rec_n : A.rec_cap;
  const int lane = threadIdx.x & 63;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t k = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; k < n; k += nw) {
    const uint32_t g = V.ilist[k];
    const int32_t j = wave_job_of(A.jobs, A.n_jobs, g, lane);
    const RSDev& R = A.rs[A.jobs[j].rs];
    const Col c = load_col(A.pile, A.G, g);
    int32_t pct = 0;
    const int need = indel_kind_needed(c, A.cf, &pct);
    int64_t win = -1;
    if (need) {
      const uint32_t key = (g << 1) | (uint32_t)(need - 1);
      const int64_t a = wave_lower_bound(A.rkey, nrec, key, lane);
      int64_t b = a;
      for (;;) {
        const int64_t t = b + lane;
        const int cnt = __popcll(__ballot(t < (int64_t)nrec && A.rkey[t] == key));
        b += cnt;
        if (cnt < 64) break;
      }
      win = wave_majority(R, A.rec, A.rval, a, b, lane).rec;
    }
    if (lane == 0) V.irec[k] = win;
  }
}

// the length pass (WRITE false); the write pass is k_vcf_write
template <bool WRITE>
__global__ void __launch_bounds__(256) k_vcf_lines(VcfArgs V) {
  const WinDev win = V.A.wins[blockIdx.x];
  const JobDev& J = V.A.jobs[win.job];
  const RSDev& R = V.A.rs[J.rs];
  const char* name = V.names + V.name_off[win.job];
  const int64_t name_len = V.name_off[win.job + 1] - V.name_off[win.job];
  const int32_t minDepth = V.A.st[win.job].min_depth;
  for (int p = win.p0 + threadIdx.x; p < win.p1; p += blockDim.x) {
    const int64_t g = J.out + p;
    if (WRITE) vcf_locus<true>(V.text + V.off[g], V, J, R, name, name_len, minDepth, g, p);
    else V.len[g] = vcf_locus<false>(nullptr, V, J, R, name, name_len, minDepth, g, p);
  }
}

// The write pass.  A chunk of 256 loci (one per thread) is formatted into
// LDS at the loci's relative offsets, shifted so that LDS and output
// addresses agree mod 4; the chunk's bytes, contiguous in the output, then
// leave LDS as aligned dword loads and stores of consecutive lanes.  A chunk
// whose text exceeds the stage is written straight to HBM.
#define PL_VCF_STAGE 57344
__global__ void __launch_bounds__(256) k_vcf_write(VcfArgs V) {
  __shared__ __attribute__((aligned(16))) char stage[PL_VCF_STAGE];
  const WinDev win = V.A.wins[blockIdx.x];
  const JobDev& J = V.A.jobs[win.job];
  const RSDev& R = V.A.rs[J.rs];
  const char* name = V.names + V.name_off[win.job];
  const int64_t name_len = V.name_off[win.job + 1] - V.name_off[win.job];
  const int32_t minDepth = V.A.st[win.job].min_depth;
  const int tid = threadIdx.x;
  for (int c0 = win.p0; c0 < win.p1; c0 += 256) {
    const int c1 = c0 + 256 < win.p1 ? c0 + 256 : win.p1;
    const int p = c0 + tid;
    const int64_t g = J.out + p;
    const uint64_t base = V.off[J.out + c0];
    const uint64_t S = V.off[J.out + c1] - base;
    const uint32_t sh = (uint32_t)(((uintptr_t)V.text + base) & 3u);   // output alignment of the chunk
    if (S + sh > PL_VCF_STAGE) {
      if (p < c1) vcf_locus<true>(V.text + V.off[g], V, J, R, name, name_len, minDepth, g, p);
      continue;
    }
    char* st = stage + sh;                                    // st and the output agree mod 4
    if (p < c1) vcf_locus<true>(st + (V.off[g] - base), V, J, R, name, name_len, minDepth, g, p);
    __syncthreads();
    char* dst = V.text + base;
    const uint32_t h0 = (4u - sh) & 3u;
    const uint32_t head = h0 < S ? h0 : (uint32_t)S;
    if ((uint32_t)tid < head) dst[tid] = st[tid];
    const uint32_t nw = (uint32_t)((S - head) >> 2);
    uint32_t* dw = (uint32_t*)(dst + head);                   // 4-byte aligned
    const uint32_t* sw = (const uint32_t*)(st + head);        // 4-byte aligned in LDS
    for (uint32_t w = tid; w < nw; w += 256) dw[w] = sw[w];
    const uint32_t t0 = head + 4 * nw;
    if ((uint64_t)(t0 + tid) < S) dst[t0 + tid] = st[t0 + tid];   // at most 3 tail bytes
    __syncthreads();
  }
}

// exclusive scan of n u32 lengths into u64 offsets (out has n + 1 entries)
__global__ void __launch_bounds__(256) k_scan64_tiles(const uint32_t* in, int64_t n, uint64_t* tile_sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint64_t s = 0;
  for (int i = threadIdx.x; i < SCAN_TILE; i += 256) {
    const int64_t k = base + i;
    if (k < n) s += in[k];
  }
  s = wave_sum(s);
  __shared__ uint64_t ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ void __launch_bounds__(1024) k_scan64_top(uint64_t* tile_sums, int64_t nt) {
  __shared__ uint64_t ws[16];
  __shared__ uint64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nt; b0 += 1024) {
    const int64_t k = b0 + threadIdx.x;
    const uint64_t v = k < nt ? tile_sums[k] : 0;
    const uint64_t inc = wave_incl_scan(v);
    if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint64_t add = carry;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) add += ws[w];
    if (k < nt) tile_sums[k] = add + inc - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = add + inc;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_scan64_apply(const uint32_t* in, int64_t n, const uint64_t* tile_sums,
                                                      uint64_t* out) {
  constexpr int PER = SCAN_TILE / 256;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * PER;
  uint32_t v[PER];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    s += v[i];
  }
  const uint64_t inc = wave_incl_scan(s);
  __shared__ uint64_t ws[4];
  if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
  __syncthreads();
  uint64_t acc = tile_sums[blockIdx.x] + inc - s;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) acc += ws[w];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    if (base + i < n) out[base + i] = acc;
    if (base + i == n - 1) out[n] = acc + v[i];
    acc += v[i];
  }
  if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

template __global__ void k_vcf_lines<false>(VcfArgs);
