// pl_track_kernels.hip — the per-locus tracks of Tracks.standardTracks
// (Tracks.scala:33-56) with Pass 2 of GenomeRegion.postProcess
// (GenomeRegion.scala:273-283) for a batch's regions, on the device.
// Included at the end of pl_kernels.hip.
//
//   k_gcv_carry/top/fill  once per contig: the value each contig locus
//                 leaves in computeGc's 100-slot circular buffer (1 for G/C,
//                 0 for A/T, else the slot's previous value: the value of
//                 the locus 100 before, or the buffer's initial 1,0,1,0...),
//                 as a scan down each of the 100 residue columns;
//   k_trk_frag    fragment coverage per locus (processBam's depth before
//                 Pass 1 added a called deletion's count to the loci under
//                 it) and its sum per region;
//   k_trk_values  the 13 values per locus: copy number from the smoothed
//                 fragment coverage over the region's mean, GC as a window
//                 sum of the buffer values (both from prefix sums staged in
//                 LDS per window), the rest from the pileup;
//   k_wig_lines   one track's text (length pass, then write pass).

// issue bits of a locus (GenomeRegion.scala:696-720), for Pilon.bed
#define PL_ISSUE_UNCONFIRMED 0x01  // !confirmed(i)
#define PL_ISSUE_CHANGED     0x02  // changed(i) || deleted(i)
#define PL_ISSUE_LOWCOV      0x04  // lowCoverage(i)
#define PL_ISSUE_HIGHCN      0x08  // copyNumber(i) > 1
#define PL_ISSUE_N           0x10  // refBase(locus(i)) == 'N'
#define PL_ISSUE_BREAK       0x20  // breakp(i)
#define PL_ISSUE_DELTA       0x40  // deltaFraction(i) >= 0.5

#define PL_GCV_MIN_ROWS 64     // contig rows (of 100 loci) per column-scan chunk, at least
#define PL_GCV_MAX_CHUNKS 2048 // chunks per contig, at most (the serial pass between)
#define PL_GC_STAGE (2048 + 149)     // GC buffer values a window needs, at most
#define PL_FRAG_STAGE (2048 + 201)   // fragment coverage a window needs, at most

struct TrackArgs {
  EngineArgs A;
  int32_t* frag;               // [G] fragment coverage
  const int32_t* jdepth;       // [G] depth of "jumps" reads left out of it (null: none)
  int64_t* jsum;               // [n_jobs] sum of frag per region
  const uint8_t* const* gcv;   // [n_jobs] the region's contig buffer values
  int32_t* vals;               // [PL_NTRACKS][G]
  // Pilon.bed: per-locus issue bits (null: not asked) and per-region sums
  uint8_t* mask;               // [G] PL_ISSUE_* bits
  int64_t* jbad;               // [n_jobs] sum of badCoverage
  int64_t* jgood;              // [n_jobs] sum of coverage
  // one track's wig text
  int32_t track;
  const char* names;           // region (contig) names, concatenated
  const int64_t* name_off;     // [n_jobs + 1]
  uint32_t* len;               // [G] bytes per locus
  const uint64_t* off;         // [G + 1]
  char* text;
};

namespace {

// 1 for G/C, 0 for A/T, 2 for any other byte (GenomeRegion.scala:173-179)
__device__ __forceinline__ uint32_t gc_class(uint8_t c) {
  return (c == 'G' || c == 'C') ? 1u : (c == 'A' || c == 'T') ? 0u : 2u;
}

// java.lang.Math.round(double) for finite arguments within Long range
__device__ __forceinline__ int64_t java_round_dev(double x) {
  const double f = floor(x);
  return (int64_t)f + ((x - f) >= 0.5 ? 1 : 0);
}

// GenomeRegion.smooth(fragCoverage, 200)(r) (GenomeRegion.scala:190-210)
// in closed form, Int arithmetic: result(r) for 101 <= r <= n-101 is
// (input(0) + sum input(r-99 .. r+100) + 100) / 200 (the first subtraction
// drops input(1), so input(0) stays in the sum); result(0..100) is never
// written (0), result(n-100 .. n-1) repeat result(n-101); n <= 200 gives
// sum / n everywhere.  F holds prefix sums (mod 2^32) of the region's
// fragment coverage from index flo on; total is the region's sum.
__device__ __forceinline__ int32_t smooth_at(const uint32_t* F, int32_t flo, int32_t n, int32_t in0, int64_t total,
                                             int32_t r) {
  if (n <= 200) return j_div((int32_t)(uint32_t)total, n);
  if (r >= n - 100) r = n - 101;
  if (r <= 100) return 0;
  const int32_t acc = j_add(in0, (int32_t)(F[r + 100 - flo + 1] - F[r - 99 - flo]));
  return j_div(j_add(acc, 100), 200);
}

template <bool WRITE>
struct WOut {
  char* p;
  uint32_t n;
  __device__ void c(char x) {
    if (WRITE) p[n] = x;
    n++;
  }
  __device__ void s(const char* t) {
    for (; *t; t++) c(*t);
  }
  __device__ void i(int32_t v) {                           // Integer.toString
    uint32_t u = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
    if (v < 0) c('-');
    uint64_t packed = 0;
    int nd = 0;
    do {
      packed |= (uint64_t)(u % 10u) << (4 * nd);
      u /= 10u;
      nd++;
    } while (u);
    for (int k = nd - 1; k >= 0; k--) c((char)('0' + ((packed >> (4 * k)) & 15u)));
  }
};

// one locus of a wig track, preceded at region index 0 by the region's
// fixedStep line (Tracks.makeTrack, Tracks.scala:168-176)
template <bool WRITE>
__device__ uint32_t wig_locus(char* out, const TrackArgs& T, int32_t job, const JobDev& J, int64_t g, int32_t r) {
  WOut<WRITE> o;
  o.p = out;
  o.n = 0;
  if (r == 0) {
    o.s("fixedStep chrom=");
    for (int64_t k = T.name_off[job]; k < T.name_off[job + 1]; k++) o.c(T.names[k]);
    o.s(" start=");
    o.i(J.start);
    o.s(" step=1\n");
  }
  o.i(T.vals[(int64_t)T.track * T.A.G + g]);
  o.c('\n');
  return o.n;
}

}  // namespace

// ---------------------------------------------------------------- GC buffer
// Locus k of a contig is column k % 100, row k / 100; a contig's rows are
// cut into chunks of `rpc` rows.  The contigs still without buffer values
// are built together: one thread per (contig, chunk, column) in the carry
// and fill passes (adjacent threads read adjacent loci), one per (contig,
// column) in the pass between them.
struct GcvJob {
  const uint8_t* ref;
  uint8_t* gcv;
  int64_t n;                 // contig length
  int64_t rpc;               // rows per chunk
  int64_t chunks;
  int64_t t0;                // first thread (chunk * 100 + column) of this contig
};

__device__ __forceinline__ int32_t gcv_job_of(const GcvJob* J, int32_t nj, int64_t t) {
  int32_t lo = 0, hi = nj - 1;
  while (lo < hi) {
    const int32_t m = (lo + hi + 1) >> 1;
    if (J[m].t0 <= t) lo = m; else hi = m - 1;
  }
  return lo;
}

// carry = the chunk's last defined value in the column (2: none)
__global__ void __launch_bounds__(256) k_gcv_carry(const GcvJob* J, int32_t nj, int64_t nt, uint8_t* carry) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const GcvJob& g = J[gcv_job_of(J, nj, t)];
  const int64_t u = t - g.t0, ch = u / 100, col = u % 100;
  const int64_t rows = (g.n + 99) / 100;
  const int64_t r0 = ch * g.rpc, r1 = r0 + g.rpc < rows ? r0 + g.rpc : rows;
  uint32_t v = 2;
  for (int64_t r = r1 - 1; r >= r0 && v == 2; r--) {
    const int64_t k = r * 100 + col;
    if (k < g.n) v = gc_class(g.ref[k]);
  }
  carry[t] = (uint8_t)v;
}

// one thread per (contig, column): the value entering each chunk, starting
// from the buffer's initial value (slot `col` holds 1 when col is even)
__global__ void __launch_bounds__(256) k_gcv_top(const GcvJob* J, int32_t nj, uint8_t* carry) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nj * 100) return;
  const GcvJob& g = J[t / 100];
  const int64_t col = t % 100;
  uint8_t run = (col & 1) == 0;
  for (int64_t ch = 0; ch < g.chunks; ch++) {
    const int64_t i = g.t0 + ch * 100 + col;
    const uint8_t c = carry[i];
    carry[i] = run;
    if (c != 2) run = c;
  }
}

__global__ void __launch_bounds__(256) k_gcv_fill(const GcvJob* J, int32_t nj, int64_t nt, const uint8_t* carry) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const GcvJob& g = J[gcv_job_of(J, nj, t)];
  const int64_t u = t - g.t0, ch = u / 100, col = u % 100;
  const int64_t rows = (g.n + 99) / 100;
  const int64_t r0 = ch * g.rpc, r1 = r0 + g.rpc < rows ? r0 + g.rpc : rows;
  const uint8_t* __restrict__ ref = g.ref;
  uint8_t* __restrict__ out = g.gcv;
  uint32_t run = carry[t];
  // eight loads in flight before the dependent selects
  for (int64_t r = r0; r < r1; r += 8) {
    uint8_t c[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int64_t k = (r + q) * 100 + col;
      c[q] = (r + q < r1 && k < g.n) ? ref[k] : (uint8_t)0;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int64_t k = (r + q) * 100 + col;
      if (r + q < r1 && k < g.n) {
        const uint32_t v = gc_class(c[q]);
        if (v != 2) run = v;
        out[k] = (uint8_t)run;
      }
    }
  }
}

// ------------------------------------------------------------ fragment cov
__global__ void __launch_bounds__(256) k_trk_frag(TrackArgs T) {
  const EngineArgs& A = T.A;
  const WinDev win = A.wins[blockIdx.x];
  const JobDev& J = A.jobs[win.job];
  const int64_t G = A.G;
  const bool reads = A.st[win.job].read_count != 0;
  int64_t s = 0, sgood = 0, sbad = 0;
  for (int p = win.p0 + threadIdx.x; p < win.p1; p += blockDim.x) {
    const int64_t g = J.out + p;
    int64_t d = 0;
    for (int b = 0; b < 4; b++) d += (uint32_t)A.pile[(int64_t)(PL_F_COUNT_A + b) * G + g];
    int32_t fc = (int32_t)(d + A.pile[(int64_t)PL_F_DELETIONS * G + g]);
    if (A.calls[g].flags & PL_CF_DELETED) {
      // the deletion's caller: the nearest locus to the left that Pass 1 did
      // not delete (GenomeRegion.scala:259-264); its count was added here
      int64_t k = g - 1;
      while (k > J.out && (A.calls[k].flags & PL_CF_DELETED)) k--;
      fc = j_sub(fc, A.pile[(int64_t)PL_F_DELETIONS * G + k]);
    }
    if (T.jdepth) fc = j_sub(fc, T.jdepth[g]);          // processBam skips jumps BAMs (GenomeRegion.scala:291-298)
    T.frag[g] = fc;
    s += fc;
    if (T.mask && reads) {                                   // pctBadOverall's Long sums (GenomeRegion.scala:99-113)
      sgood += (int32_t)(d + A.pile[(int64_t)PL_F_DELETIONS * G + g]);
      sbad += A.pile[(int64_t)PL_F_BADPAIR * G + g];
    }
  }
  s = wave_sum(s);
  __shared__ int64_t ws[3][4];
  if ((threadIdx.x & 63) == 0) ws[0][threadIdx.x >> 6] = s;
  if (T.mask) {
    sgood = wave_sum(sgood);
    sbad = wave_sum(sbad);
    if ((threadIdx.x & 63) == 0) { ws[1][threadIdx.x >> 6] = sgood; ws[2][threadIdx.x >> 6] = sbad; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t t = ws[0][0] + ws[0][1] + ws[0][2] + ws[0][3];
    if (t) atomicAdd((unsigned long long*)&T.jsum[win.job], (unsigned long long)t);
    if (T.mask) {
      const int64_t tg = ws[1][0] + ws[1][1] + ws[1][2] + ws[1][3], tb = ws[2][0] + ws[2][1] + ws[2][2] + ws[2][3];
      if (tg) atomicAdd((unsigned long long*)&T.jgood[win.job], (unsigned long long)tg);
      if (tb) atomicAdd((unsigned long long*)&T.jbad[win.job], (unsigned long long)tb);
    }
  }
}

// ------------------------------------------------------------ track values
__global__ void __launch_bounds__(256) k_trk_values(TrackArgs T) {
  __shared__ int32_t P[PL_GC_STAGE + 1];
  __shared__ uint32_t F[PL_FRAG_STAGE + 1];
  __shared__ int32_t ws[4];
  __shared__ uint32_t wf[4];
  const EngineArgs& A = T.A;
  const WinDev win = A.wins[blockIdx.x];
  const JobDev& J = A.jobs[win.job];
  const JobStat& S = A.st[win.job];
  const RSDev& R = A.rs[J.rs];
  const uint8_t* gcv = T.gcv[win.job];
  const int64_t n = R.ref_len;
  const int64_t G = A.G;

  // computeGc (GenomeRegion.scala:160-188) tests and indexes the 0-based
  // contig loci with the region's 1-based bounds, so region index p holds
  // the count for center c = start + p (0-based): the buffer after locus
  // c + 50 (after locus c itself in the contig's last 50 loci; none when
  // c is past the contig).  The counts need buffer values at loci
  // [start + p0 - 99, start + p1 + 50); their prefix sums go to LDS.
  const int64_t base = (int64_t)J.start + win.p0 - 99;
  const int32_t SN = (win.p1 - win.p0) + 149;
  {
    const int32_t per = (SN + 255) / 256;
    const int32_t i0 = threadIdx.x * per, i1 = i0 + per < SN ? i0 + per : SN;
    int32_t s = 0;
    for (int32_t i = i0; i < i1; i++) {
      const int64_t k = base + i;
      s += k < 0 ? (int32_t)((k & 1) == 0) : k < n ? (int32_t)gcv[k] : 0;
    }
    const int32_t inc = wave_incl_scan(s);
    if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
    __syncthreads();
    int32_t acc = inc - s;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) acc += ws[w];
    if (threadIdx.x == 0) P[0] = 0;
    for (int32_t i = i0; i < i1; i++) {
      const int64_t k = base + i;
      acc += k < 0 ? (int32_t)((k & 1) == 0) : k < n ? (int32_t)gcv[k] : 0;
      P[i + 1] = acc;
    }
  }
  // Fragment coverage of region indices [flo, fhi), as LDS prefix sums: the
  // smoothing window (r-99 .. r+100, r clamped to n-101 at the right edge)
  // and the +-100 neighbours of delta/dip of every locus of the window.
  const int32_t nreg = J.size;
  const int32_t edge = nreg - 101 > 0 ? nreg - 101 : 0;
  const int32_t flo0 = (win.p0 < edge ? win.p0 : edge) - 100;
  const int32_t flo = flo0 > 0 ? flo0 : 0;
  const int32_t fhi = nreg < win.p1 + 101 ? nreg : win.p1 + 101;
  {
    const int32_t FN = fhi - flo;
    const int32_t per = (FN + 255) / 256;
    const int32_t i0 = threadIdx.x * per, i1 = i0 + per < FN ? i0 + per : FN;
    uint32_t s = 0;
    for (int32_t i = i0; i < i1; i++) s += (uint32_t)T.frag[J.out + flo + i];
    const uint32_t inc = wave_incl_scan(s);
    if ((threadIdx.x & 63) == 63) wf[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t acc = inc - s;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) acc += wf[w];
    if (threadIdx.x == 0) F[0] = 0;
    for (int32_t i = i0; i < i1; i++) {
      acc += (uint32_t)T.frag[J.out + flo + i];
      F[i + 1] = acc;
    }
    __syncthreads();
  }
  const int32_t in0 = T.frag[J.out];

  const bool reads = S.read_count != 0;
  const double baseCov = (double)T.jsum[win.job] / (double)J.size;       // NormalDistribution.mean
  int64_t pbo = 0;                                                       // pctBadOverall, Long
  if (T.mask) pbo = round_div_l(100 * T.jbad[win.job], T.jgood[win.job] + T.jbad[win.job]);
  for (int p = win.p0 + threadIdx.x; p < win.p1; p += blockDim.x) {
    const int64_t g = J.out + p;
    // the pileup fields the tracks read (10 of the 18)
    int64_t cnt = 0, qsum = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      cnt += (uint32_t)A.pile[(int64_t)(PL_F_COUNT_A + b) * G + g];
      qsum += A.pile[(int64_t)(PL_F_QUAL_A + b) * G + g];
    }
    const int32_t mqSum = A.pile[(int64_t)PL_F_MQSUM * G + g];
    const int32_t qSum = A.pile[(int64_t)PL_F_QSUM * G + g];
    const uint16_t f = A.calls[g].flags;
    if (f & PL_CF_WIDE) {          // the exact Long qualSums of a wide column
      Col c;
      wide_patch(c, A.wide, A.wide_cap, A.win_wide, J, A.W, g);
      qsum = c.qs[0] + c.qs[1] + c.qs[2] + c.qs[3];
    }
    const int64_t dep = cnt + A.pile[(int64_t)PL_F_DELETIONS * G + g];
    // the final disposition (as k_final_flags)
    int disp = (f & PL_CF_DISP_MASK) >> PL_CF_DISP_SHIFT;
    if (!reads || (f & PL_CF_DELETED) || dep < S.min_depth) disp = PL_D_NONE;
    // processBam leaves the pileup tracks at 0 in a region without reads
    const int32_t cov = reads ? (int32_t)dep : 0;
    const int32_t bad = reads ? A.pile[(int64_t)PL_F_BADPAIR * G + g] : 0;
    int16_t cn = 0;
    if (reads && baseCov > 0)
      cn = (int16_t)java_round_dev((double)smooth_at(F, flo, nreg, in0, T.jsum[win.job], p) / baseCov);
    // deltaCoverage / dipCoverage at radius 100 (GenomeRegion.scala:672-693)
    const int32_t il = (p > 100 ? p - 100 : 0) - flo, ir = (p + 100 < nreg - 1 ? p + 100 : nreg - 1) - flo;
    const int32_t l = (int32_t)(F[il + 1] - F[il]);
    const int32_t rr = (int32_t)(F[ir + 1] - F[ir]);
    const int32_t ce = (int32_t)(F[p - flo + 1] - F[p - flo]);
    const int32_t d = j_sub(l, rr);
    const int64_t cc = (int64_t)J.start + p;
    int32_t gc = 0;
    if (cc < n) {
      const int64_t L = cc >= n - 50 ? cc : cc + 50;
      gc = (int32_t)(int8_t)(P[L - base + 1] - P[L - 99 - base]);
    }
    int32_t v[PL_NTRACKS];
    v[0] = (disp == PL_D_SNP || disp == PL_D_INS || disp == PL_D_DEL) ? 1 : 0;   // Changes
    v[1] = disp == PL_D_CONFIRMED ? 0 : 1;                                       // Unconfirmed
    v[2] = (int32_t)cn - 1;                                                      // Copy Number
    v[3] = cov;                                                                  // Coverage
    v[4] = bad;                                                                  // Bad Coverage
    v[5] = j_add(cov, bad) > 0 ? j_div(j_mul(bad, 100), j_add(cov, bad)) : 0;   // Pct Bad
    v[6] = d < 0 ? j_sub(0, d) : d;                                              // Delta Coverage
    v[7] = j_add(j_sub(l, ce), j_sub(rr, ce));                                   // Dip Coverage
    v[8] = reads ? A.pile[(int64_t)PL_F_PHYSCOV * G + g] : 0;                    // Physical Coverage
    v[9] = reads ? (int32_t)(int16_t)A.pile[(int64_t)PL_F_CLIPS * G + g] : 0;    // Clipped Alignments
    v[10] = reads ? (int32_t)(int8_t)round_div_l(qsum, (int64_t)mqSum) : 0;   // Weighted Qual (PileUp.scala:60-62)
    v[11] = reads ? (int32_t)(int8_t)round_div_l(qsum, (int64_t)qSum) : 0;    // Weighted MQ (PileUp.scala:56-58)
    v[12] = gc;                                                                  // GC
#pragma unroll
    for (int t = 0; t < PL_NTRACKS; t++) T.vals[(int64_t)t * G + g] = v[t];
    if (T.mask) {
      // the loci tests behind Pilon.bed's regions (GenomeRegion.scala:696-720)
      const int32_t mmd = A.cf.min_min_depth;
      uint8_t rb = R.ref[(int64_t)J.start + p - 1];
      if (rb >= 'a' && rb <= 'z') rb = (uint8_t)(rb - 32);
      const bool isN = rb == 'N';
      const bool low = cov < mmd && !isN;
      const bool clip = cov >= mmd && pct_i(v[9], cov) >= 33;
      const bool bad_ = (int64_t)pct_i(bad, j_add(cov, bad)) > pbo + 20;
      const bool dip = (double)v[7] / baseCov >= 1.5;
      uint8_t m = 0;
      if (disp != PL_D_CONFIRMED) m |= PL_ISSUE_UNCONFIRMED;
      if (v[0] || (f & PL_CF_DELETED)) m |= PL_ISSUE_CHANGED;
      if (low) m |= PL_ISSUE_LOWCOV;
      if (cn > 1) m |= PL_ISSUE_HIGHCN;
      if (isN) m |= PL_ISSUE_N;
      if (low || clip || bad_ || dip) m |= PL_ISSUE_BREAK;
      if ((double)v[6] / baseCov >= 0.5) m |= PL_ISSUE_DELTA;
      T.mask[g] = m;
    }
  }
}

template <bool WRITE>
__global__ void __launch_bounds__(256) k_wig_lines(TrackArgs T) {
  const WinDev win = T.A.wins[blockIdx.x];
  const JobDev& J = T.A.jobs[win.job];
  for (int p = win.p0 + threadIdx.x; p < win.p1; p += blockDim.x) {
    const int64_t g = J.out + p;
    if (WRITE) wig_locus<true>(T.text + T.off[g], T, win.job, J, g, p);
    else T.len[g] = wig_locus<false>(nullptr, T, win.job, J, g, p);
  }
}

template __global__ void k_wig_lines<false>(TrackArgs);
template __global__ void k_wig_lines<true>(TrackArgs);
