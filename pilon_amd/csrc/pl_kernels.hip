// pl_kernels.hip — gfx950 kernels of the pileup path.
//
//   k_prep           per read, once per resident read set: alignment end,
//                    validity/adjMq, touched span, fragment, outliers
//   k_scan_*         exclusive scans (keep-read prefix, radix histograms)
//   k_plan           per window: the sorted read range it must scan
//   k_jobprep        per region: readCount (BamFile.process), P ops, errors
//   k_window         the hot kernel: one workgroup owns a window of W
//                    positions, accumulates every counter in LDS, scans
//                    physCov, and writes counters + BaseCall once
//   k_rs_*           LSD radix sort (evidence records, change candidates)
//   k_indel_pass     BaseCall of columns whose indel strings matter
//   k_finalize       Pass 1 per region: minDepth, deletion pass
//   k_recompute      BaseCall of columns changed by the deletion pass
//   k_final_flags    final dispositions for export
//
// Reference: broadinstitute/pilon v1.24, src/main/scala/org/broadinstitute/pilon/
#include <hip/hip_runtime.h>
#include "pl_core.h"
#include "pl_kernels.h"

namespace {

__device__ __forceinline__ int64_t lower_pos_dev(const int32_t* pos, int64_t lo, int64_t hi, int64_t x) {
  while (lo < hi) {
    int64_t m = (lo + hi) >> 1;
    if ((int64_t)pos[m] < x) lo = m + 1; else hi = m;
  }
  return lo;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane ? ((~0ull) >> (64 - lane)) : 0ull;
}

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// slot of this lane in a list grown by the lanes of the wave whose pred is
// set: one atomic on the list counter per wave
__device__ __forceinline__ uint32_t wave_append(uint32_t* ctr, bool pred) {
  const uint64_t m = __ballot(pred);
  if (!m) return 0u;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0u;
  if ((int)(threadIdx.x & 63) == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
  base = __shfl(base, leader, 64);
  return base + (uint32_t)__popcll(m & lanemask_lt());
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Wave-wide integer scans through DPP lane moves (no LDS round trips): a
// Hillis-Steele scan inside each row of 16 lanes (row_shr 1, 2, 4, 8, lanes
// shifted in from outside the row read 0), then the last lane of row 0 added
// into row 1 and of row 2 into row 3 (row_bcast:15), and the last lane of
// row 1 into rows 2 and 3 (row_bcast:31).  Integer adds only, so the result
// equals the shuffle scan's; all 64 lanes must be active.
template <int CTRL, int ROW = 0xf>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW, 0xf, true);
}
template <class T>
__device__ __forceinline__ T wave_scan_dpp(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit integers");
  if constexpr (sizeof(T) == 4) {
    uint32_t x = (uint32_t)v;
    x += dpp_move<0x111>(x);
    x += dpp_move<0x112>(x);
    x += dpp_move<0x114>(x);
    x += dpp_move<0x118>(x);
    x += dpp_move<0x142, 0xa>(x);
    x += dpp_move<0x143, 0xc>(x);
    return (T)x;
  } else {
    uint64_t x = (uint64_t)v;
    auto mv = [](auto ctrl, auto row, uint64_t y) {
      return (uint64_t)dpp_move<decltype(ctrl)::value, decltype(row)::value>((uint32_t)y) |
             ((uint64_t)dpp_move<decltype(ctrl)::value, decltype(row)::value>((uint32_t)(y >> 32)) << 32);
    };
    using std::integral_constant;
    x += mv(integral_constant<int, 0x111>{}, integral_constant<int, 0xf>{}, x);
    x += mv(integral_constant<int, 0x112>{}, integral_constant<int, 0xf>{}, x);
    x += mv(integral_constant<int, 0x114>{}, integral_constant<int, 0xf>{}, x);
    x += mv(integral_constant<int, 0x118>{}, integral_constant<int, 0xf>{}, x);
    x += mv(integral_constant<int, 0x142>{}, integral_constant<int, 0xa>{}, x);
    x += mv(integral_constant<int, 0x143>{}, integral_constant<int, 0xc>{}, x);
    return (T)x;
  }
}
// lane 63's value in every lane (wave-uniform)
template <class T>
__device__ __forceinline__ T lane63(T v) {
  if constexpr (sizeof(T) == 4) return (T)__builtin_amdgcn_readlane((int)v, 63);
  else
    return (T)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), 63) << 32) |
               (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(uint64_t)v, 63));
}
template <class T>
__device__ __forceinline__ T wave_total_dpp(T v) { return lane63(wave_scan_dpp(v)); }

}  // namespace

// ------------------------------------------------------------------ k_prep
// segments are built for multi-op reads (not RI_SIMPLE) that addRead processes
PL_HD bool wants_segs(uint32_t info) {
  return (info & RI_KEEP) && !(info & (RI_SIMPLE | RI_EXOTIC | RI_BAD | RI_LR_MASK));
}

// the read set's scalars (scal[0], [1], [2], [6]: maxima, -1 = none;
// flags: scal[4], scal[5]) gathered over a thread's reads, then over its
// wave, so one wave gives one atomic per scalar
struct PrepAcc {
  int32_t back = -1, lead = -1, span = -1, fp = -1;
  uint32_t flags = 0;
};
__device__ __forceinline__ void prep_flush(PrepAcc a, int32_t* scal) {
  for (int d = 32; d >= 1; d >>= 1) {
    a.back = max(a.back, __shfl_xor(a.back, d, 64));
    a.lead = max(a.lead, __shfl_xor(a.lead, d, 64));
    a.span = max(a.span, __shfl_xor(a.span, d, 64));
    a.fp = max(a.fp, __shfl_xor(a.fp, d, 64));
    a.flags |= (uint32_t)__shfl_xor((int)a.flags, d, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (a.back >= 0) atomicMax(&scal[0], a.back);
    if (a.lead >= 0) atomicMax(&scal[1], a.lead);
    if (a.span >= 0) atomicMax(&scal[2], a.span);
    if (a.fp >= 0) atomicMax(&scal[6], a.fp);
    if (a.flags & 1u) atomicOr(&scal[4], 1);
    if (a.flags & 2u) atomicOr(&scal[5], 1);
  }
}

// k_prep's work on read r of set R
__device__ __forceinline__ void prep_one(const RSDev& R, int64_t r, const KConf& cf, uint32_t* keepflag,
                                         uint32_t* segcount, int32_t* scal, int64_t* outliers, int64_t out_cap,
                                         PrepAcc& acc) {
  PrepOut o = prep_read(R, r, cf);
  RMeta m = o.m;
  // a read without QUAL piles defaultQual at every base (PileUpRegion.scala:114-115):
  // written into its quality bytes once, so the window kernel reads quals unconditionally
  if (m.info & RI_NOQUAL) {
    int8_t* qd = (int8_t*)R.quals + R.seq_off[r];
    const int64_t L = R.seq_off[r + 1] - R.seq_off[r];
    for (int64_t k = 0; k < L; k++) qd[k] = (int8_t)cf.default_qual;
  }
  segcount[r] = wants_segs(m.info) ? (uint32_t)build_segs(R, r, m.info, cf, nullptr) : 0u;
  const bool keep = (m.info & RI_KEEP) != 0;
  if (keep) {
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    if (m.fp_lo <= m.fp_hi) { lo = m.fp_lo; hi = m.fp_hi; }
    if (m.fr_lo < m.fr_hi) {
      lo = lo < m.fr_lo ? lo : m.fr_lo;
      hi = hi > (int64_t)m.fr_hi - 1 ? hi : (int64_t)m.fr_hi - 1;
    }
    if (lo <= hi) {
      const int64_t lead = m.pos - lo, back = hi - m.pos;
      if (lead > PL_MAX_LEAD || back > PL_MAX_BACK || R.seq_off[r + 1] - R.seq_off[r] >= PL_MAX_SEQ) {
        m.info |= RI_OUTLIER;
        int k = atomicAdd(&scal[3], 1);
        if (k < out_cap) outliers[k] = r;
      } else {
        acc.back = max(acc.back, (int32_t)(back > 0 ? back : 0));
        acc.lead = max(acc.lead, (int32_t)(lead > 0 ? lead : 0));
        if (m.fp_lo <= m.fp_hi && m.fp_hi > m.pos) acc.fp = max(acc.fp, m.fp_hi - m.pos);
      }
    }
    const int32_t spanEnd = m.aend != 0 ? m.aend : m.pos;
    acc.span = max(acc.span, spanEnd - m.pos > 0 ? spanEnd - m.pos : 0);
    if (m.info & RI_UNK_MASK) acc.flags |= 1u;
    if (m.info & (RI_BAD | RI_EDGE_CLIP)) acc.flags |= 2u;
  }
  R.meta[r] = m;
  R.ftail[r] = frag_tail(m);
  keepflag[r] = keep ? 1u : 0u;
}

__global__ void k_prep(RSDev R, KConf cf, uint32_t* keepflag, uint32_t* segcount, int32_t* scal, int64_t* outliers,
                       int64_t out_cap) {
  const int64_t n = R.n;
  PrepAcc acc;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    prep_one(R, r, cf, keepflag, segcount, scal, outliers, out_cap, acc);
  prep_flush(acc, scal);
}

// the set of block b when set s has blocks [blk[s], blk[s + 1])
__device__ __forceinline__ int32_t set_of_block(const int64_t* blk, int32_t n_sets, int64_t b) {
  int32_t lo = 0, hi = n_sets;
  while (hi - lo > 1) {
    const int32_t mid = (lo + hi) / 2;
    if (blk[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

// k_prep over many sets in one launch: set s's reads on blocks [blk[s],
// blk[s + 1]), its flags at keep + at[s], its outlier list outl[s] of
// ocap[s] entries
__global__ void k_prep_sets(const RSDev* sets, const int64_t* blk, int32_t n_sets, KConf cf, uint32_t* keep,
                            uint32_t* segc, const int64_t* at, int64_t* const* outl, const int64_t* ocap) {
  const int32_t s = set_of_block(blk, n_sets, blockIdx.x);
  const RSDev& R = sets[s];
  const int64_t nb = blk[s + 1] - blk[s];
  PrepAcc acc;
  for (int64_t r = (blockIdx.x - blk[s]) * blockDim.x + threadIdx.x; r < R.n; r += nb * blockDim.x)
    prep_one(R, r, cf, keep + at[s], segc + at[s], R.scal, outl[s], ocap[s], acc);
  prep_flush(acc, R.scal);   // every wave of a block is in set s
}

// fills the segment table; RMeta.op0 of a multi-op read becomes its first
// segment and the info word carries the count (RI_SEG_MAX = too many: walked)
__device__ __forceinline__ void segfill_one(const RSDev& R, int64_t r, const KConf& cf, const uint32_t* segoff,
                                            Seg* segs) {
  RMeta m = R.meta[r];
  if (!wants_segs(m.info)) return;
  const uint32_t o = segoff[r], c = segoff[r + 1] - o;
  if (c) build_segs(R, r, m.info, cf, segs + o);
  m.op0 = o;
  m.info = (m.info & ~RI_SEG_MASK) | ((c < RI_SEG_MAX ? c : (uint32_t)RI_SEG_MAX) << RI_SEG_SHIFT);
  R.meta[r] = m;
}

__global__ void k_segfill(RSDev R, KConf cf, const uint32_t* segoff, Seg* segs) {
  const int64_t n = R.n;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x)
    segfill_one(R, r, cf, segoff, segs);
}

// k_segfill over many sets in one launch: set s's segment offsets at
// segoff + at[s], its table segs[s]
__global__ void k_segfill_sets(const RSDev* sets, const int64_t* blk, int32_t n_sets, KConf cf, const uint32_t* segoff,
                               const int64_t* at, Seg* const* segs) {
  const int32_t s = set_of_block(blk, n_sets, blockIdx.x);
  const RSDev& R = sets[s];
  const int64_t nb = blk[s + 1] - blk[s];
  for (int64_t r = (blockIdx.x - blk[s]) * blockDim.x + threadIdx.x; r < R.n; r += nb * blockDim.x)
    segfill_one(R, r, cf, segoff + at[s], segs[s]);
}

// ------------------------------------------------------------------ scans
// Exclusive scan of n uint32 in tiles of SCAN_TILE; out has n+1 entries.
__global__ void __launch_bounds__(256) k_scan_tiles(const uint32_t* in, int64_t n, uint32_t* tile_sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
  for (int i = threadIdx.x; i < SCAN_TILE; i += 256) {
    const int64_t k = base + i;
    if (k < n) s += in[k];
  }
  s = wave_sum(s);
  __shared__ uint32_t ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* tile_sums, int64_t nt) {
  __shared__ uint32_t ws[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < nt; b0 += 1024) {
    const int64_t k = b0 + threadIdx.x;
    uint32_t v = k < nt ? tile_sums[k] : 0;
    uint32_t inc = wave_incl_scan(v);
    if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint32_t add = carry;
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) add += ws[w];
    if (k < nt) tile_sums[k] = add + inc - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = add + inc;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(256) k_scan_apply(const uint32_t* in, int64_t n, const uint32_t* tile_sums, uint32_t* out) {
  // each thread owns SCAN_TILE/256 consecutive items
  constexpr int PER = SCAN_TILE / 256;
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * PER;
  uint32_t v[PER];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < PER; i++) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    s += v[i];
  }
  uint32_t inc = wave_incl_scan(s);
  __shared__ uint32_t ws[4];
  if ((threadIdx.x & 63) == 63) ws[threadIdx.x >> 6] = inc;
  __syncthreads();
  uint32_t acc = tile_sums[blockIdx.x] + inc - s;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) acc += ws[w];
#pragma unroll
  for (int i = 0; i < PER; i++) {
    if (base + i < n) out[base + i] = acc;
    if (base + i == n - 1) out[n] = acc + v[i];
    acc += v[i];
  }
  if (n == 0 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

// ------------------------------------------------------------------ plan
__global__ void k_plan(const RSDev* rsets, const JobDev* jobs, WinDev* wins, int32_t n_wins, uint32_t* cost) {
  const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= n_wins) return;
  WinDev& win = wins[w];
  const JobDev& J = jobs[win.job];
  const RSDev& R = rsets[J.rs];
  const int64_t L0 = (int64_t)J.start + win.p0, L1 = (int64_t)J.start + win.p1;
  const int32_t back = R.scal[0], lead = R.scal[1];
  win.lo = lower_pos_dev(R.pos, 0, R.n, L0 - back);
  win.hi = lower_pos_dev(R.pos, win.lo, R.n, L1 + lead);
  win.plo = fragment_only_end(R.pos, win.lo, win.hi, J, L0, back, R.scal[6]);
  // descending-cost order key for the work queue (largest windows first):
  // 16 bits, so the sort takes two passes; windows beyond 65,535 cost units
  // share the front of the queue
  const int64_t c = (win.hi - win.plo) + (win.plo - win.lo) / 8 + (int64_t)(win.p1 - win.p0) / 16;
  cost[w] = 0xffffu - (uint32_t)(c > 0xffffll ? 0xffffll : c);
}

// the window heads in processing order (worder null: window order)
__global__ void k_win_heads(const JobDev* jobs, const WinDev* wins, const int32_t* worder, int32_t n_wins,
                            WinHead* heads) {
  const int32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_wins) return;
  const int32_t wi = worder ? worder[q] : q;
  const WinDev w = wins[wi];
  const JobDev& J = jobs[w.job];
  WinHead h;
  h.wi = wi; h.job = w.job; h.p0 = w.p0; h.p1 = w.p1;
  h.lo = w.lo; h.hi = w.hi; h.plo = w.plo; h.out = J.out;
  h.rs = J.rs; h.start = J.start; h.stop = J.stop; h.qs = J.qs; h.qe = J.qe;
  h.pad[0] = h.pad[1] = h.pad[2] = 0;
  heads[q] = h;
}

// ------------------------------------------------------------------ jobprep
// readCount of BamFile.process (BamFile.scala:118-139): kept reads whose span
// overlaps [qs, qe], from the keep-prefix plus the reads straddling qs.
__global__ void __launch_bounds__(64) k_jobprep(const RSDev* rsets, const uint32_t* const* keep_prefix,
                                                const JobDev* jobs, JobStat* st) {
  const JobDev& J = jobs[blockIdx.x];
  const RSDev& R = rsets[J.rs];
  const uint32_t* kp = keep_prefix[J.rs];
  JobStat& S = st[blockIdx.x];
  const int lane = threadIdx.x;
  const int64_t a = lower_pos_dev(R.pos, 0, R.n, J.qs);
  const int64_t b = lower_pos_dev(R.pos, a, R.n, (int64_t)J.qe + 1);
  const int64_t a1 = lower_pos_dev(R.pos, a, b, (int64_t)J.qs + 1);
  const int64_t a2 = lower_pos_dev(R.pos, 0, a, (int64_t)J.qs - R.scal[2]);
  int64_t adj = 0;
  // reads at pos == qs whose span ends before it (no reference-consuming op)
  for (int64_t r = a + lane; r < a1; r += 64) {
    const RMeta m = R.meta[r];
    if ((m.info & RI_KEEP) && m.aend != 0 && m.aend < J.qs) adj--;
  }
  // reads that start before qs and reach it
  for (int64_t r = a2 + lane; r < a; r += 64) {
    const RMeta m = R.meta[r];
    const int32_t se = m.aend != 0 ? m.aend : m.pos;
    if ((m.info & RI_KEEP) && se >= J.qs) adj++;
  }
  adj = wave_sum(adj);
  // P ops and malformed/edge reads: only scanned when the set has any
  int64_t unk = 0, bad_r = INT64_MAX, edge_r = INT64_MAX;
  if (R.scal[4] || R.scal[5]) {
    for (int64_t r = a2 + lane; r < b; r += 64) {
      const RMeta m = R.meta[r];
      if (!(m.info & RI_KEEP)) continue;
      const int32_t se = m.aend != 0 ? m.aend : m.pos;
      if (!(m.pos <= J.qe && se >= J.qs)) continue;
      unk += (m.info & RI_UNK_MASK) >> RI_UNK_SHIFT;
      if ((m.info & RI_BAD) && r < bad_r) bad_r = r;
      if ((m.info & RI_EDGE_CLIP) && r < edge_r) edge_r = r;
    }
    unk = wave_sum(unk);
    for (int d = 32; d >= 1; d >>= 1) {
      int64_t o1 = __shfl_xor(bad_r, d, 64), o2 = __shfl_xor(edge_r, d, 64);
      bad_r = o1 < bad_r ? o1 : bad_r;
      edge_r = o2 < edge_r ? o2 : edge_r;
    }
  }
  if (lane == 0) {
    S.read_count = (int64_t)kp[b] - (int64_t)kp[a] + adj;
    S.unknown_ops = unk;
    if (bad_r != INT64_MAX) atomicMin((unsigned long long*)&S.err_read_inval, (unsigned long long)bad_r);
    if (edge_r != INT64_MAX) atomicMin((unsigned long long*)&S.err_read_edge, (unsigned long long)edge_r);
  }
}

__global__ void k_init_stats(JobStat* st, int32_t n_jobs, uint32_t* counters, int n_counters) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_counters) counters[i] = 0;
  const int64_t words = (int64_t)n_jobs * (sizeof(JobStat) / 4);
  for (int64_t k = i; k < words; k += (int64_t)gridDim.x * blockDim.x) ((uint32_t*)st)[k] = 0;
}
// identity permutation + its length, for the window cost sort (a kernel,
// not a host copy, so the whole batch pipeline can be captured in a graph)
__global__ void k_iota(uint32_t* v, int32_t n, uint32_t* n_out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = (uint32_t)i;
  if (blockIdx.x == 0 && threadIdx.x == 0) *n_out = (uint32_t)n;
}
__global__ void k_init_stats2(JobStat* st, int32_t n_jobs) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n_jobs) { st[j].err_read_edge = INT64_MAX; st[j].err_read_inval = INT64_MAX; }
}

// ------------------------------------------------------------------ window
namespace {

struct LdsSink {
  int W;
  unsigned long long *bs, *mqq, *ins_, *del_;
  uint32_t *bad_, *clp_;
  int32_t *pc, *is;
  IndelRec* rec;
  uint32_t* rec_n;
  uint32_t rec_cap;
  IndelRec* lrec;        // the window's record buffer in LDS (PL_WIN_RECS)
  uint32_t* lrec_n;
  int32_t* overflow;
  JobStat* st;
  int64_t bcount, aligned_n;
  __device__ void frag(int p, int d, int32_t size) { atomicAdd(&pc[p], d); atomicAdd(&is[p], size); }
  __device__ void aligned(int n) { aligned_n += n; }
  __device__ void base(int p, int b, int32_t q, int32_t mq1) {
    atomicAdd(&bs[b * W + p], (unsigned long long)pack_base(q, mq1));
    atomicAdd(&mqq[p], (unsigned long long)pack_mq_q(mq1, q));
  }
  __device__ void count_base() { bcount++; }
  __device__ void bad(int p, int n) { atomicAdd(&bad_[p], (uint32_t)n); }
  __device__ void clip(int p) { atomicAdd(&clp_[p], 1u); }
  // records collect in LDS and reach HBM with one atomic per window; a
  // window with more than PL_WIN_RECS writes the rest directly
  __device__ void emit(const IndelRec& e) {
    const uint32_t j = atomicAdd(lrec_n, 1u);
    if (j < PL_WIN_RECS) { lrec[j] = e; return; }
    uint32_t i = atomicAdd(rec_n, 1u);
    if (i < rec_cap) rec[i] = e;
    else *overflow = 1;
  }
  __device__ void ins(int p, int32_t q, int32_t mq1, const IndelRec& e) {
    atomicAdd(&ins_[p], (unsigned long long)((1ull << 32) + (uint64_t)mq1));
    atomicAdd(&mqq[p], (unsigned long long)((uint64_t)(int64_t)q << 32));
    emit(e);
  }
  __device__ void del(int p, int32_t q, int32_t mq1, const IndelRec& e) {
    atomicAdd(&del_[p], (unsigned long long)((1ull << 32) + (uint64_t)mq1));
    atomicAdd(&mqq[p], (unsigned long long)pack_mq_q(mq1, q));
    emit(e);
  }
  __device__ void edge(int64_t r) { atomicMin((unsigned long long*)&st->err_read_edge, (unsigned long long)r); }
};

// global-address-space views: loads through them compile to global_load_*
// (vector memory counter only) instead of flat loads that also hold the LDS
// counter the atomics wait on
typedef const int8_t __attribute__((address_space(1))) gi8;
typedef const uint8_t __attribute__((address_space(1))) gu8;
typedef const uint32_t __attribute__((address_space(1))) gu32;

struct __attribute__((aligned(16))) RunL {  // one run of a chunk of reads, one 16-byte LDS read
  int32_t off;     // first chunk of the run in the chunk stream
  uint32_t pi;     // window position | EF_* flags | mq1 << 16
  int32_t q0;      // sequence offset of the run's first base, from the reads' 4-aligned base
  int32_t pad;     // run length in bases
};

__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

}  // namespace

// One workgroup owns a window of W positions; every counter of those
// positions lives in LDS until the window is done.  Reads are taken in tiles
// of BLOCK (one per thread): the headers of a tile are loaded in parallel,
// each read is classified, and the trusted in-window bases of all "simple"
// reads (one M/=/X op, 2-bit bases) are processed as one flattened stream in
// which consecutive lanes take consecutive bases — coalesced loads and
// conflict-free LDS atomics.  Reads with clips, indels or exotic bases go
// through walk_read, one wave per read.
static_assert(PL_CHUNK % 2 == 1 && PL_CHUNK <= 31, "an odd chunk of at most 31 bases (two 16-base words)");

template <int BLOCK>
__global__ void __launch_bounds__(BLOCK, PL_WIN_WAVES_EU) k_window(EngineArgs A, int persistent) {
  constexpr int NW = BLOCK / 64;
  extern __shared__ __align__(16) unsigned long long lds[];
  const int W = A.W;
  unsigned long long* bs = lds;
  unsigned long long* mqq = bs + 4 * W;
  unsigned long long* insA = mqq + W;
  unsigned long long* delA = insA + W;
  uint32_t* badA = (uint32_t*)(delA + W);
  uint32_t* clpA = badA + W;
  int32_t* pcA = (int32_t*)(clpA + W);
  int32_t* isA = pcA + W;
  int32_t* bdA = isA + W;   // badPair differences of bad runs (clips, invalid reads' M bases)
  uint32_t* hist = (uint32_t*)(bdA + W);
  __shared__ RunL s_run[NW * (2 * 64 + 1)];
  __shared__ int32_t s_w;
  __shared__ int32_t ws_a[NW], ws_b[NW], ws_c[NW];
  __shared__ unsigned long long red[2][NW];
  __shared__ int32_t s_ovf;
  __shared__ IndelRec s_rec[PL_WIN_RECS];
  __shared__ uint32_t s_nrec, s_recbase;
  // the current window's head and the next one's, fetched while the
  // current window is piled (persistent form), with their queue slots
  __shared__ WinHead s_head[2];
  __shared__ int32_t s_slot[2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lds_words = (PL_WIN_LDS_PER_POS * W) / 8 + 256 / 2;
  const KConf& cf = A.cf;
  (void)s_w;

  if (tid == 0) {
    const int32_t q = persistent ? atomicAdd(A.queue, 1) : (int32_t)blockIdx.x;
    s_slot[0] = q;
    if (q < A.n_wins) s_head[0] = A.heads[q];
  }
  __syncthreads();
  for (int iter = 0;; iter++) {
    const int cur = iter & 1;
    if (__builtin_amdgcn_readfirstlane(s_slot[cur]) >= A.n_wins) break;
    if (!persistent && iter > 0) break;
    // the next window's slot: taken now, its head loaded after the read loop
    int32_t next_slot = 0;
    if (persistent && tid == 0) next_slot = atomicAdd(A.queue, 1);
    // the head through LDS, made wave-uniform so every field is a scalar
    WinHead win;
    {
      const int32_t* hw = (const int32_t*)&s_head[cur];
      int32_t* dw = (int32_t*)&win;
#pragma unroll
      for (int k = 0; k < (int)(sizeof(WinHead) / 4); k++) dw[k] = __builtin_amdgcn_readfirstlane(hw[k]);
    }
    const int32_t wi = win.wi;
    JobDev J;
    J.rs = win.rs; J.start = win.start; J.stop = win.stop; J.qs = win.qs; J.qe = win.qe;
    J.size = win.stop - win.start + 1; J.win0 = 0; J.nwin = 0; J.out = win.out; J.rlo = 0; J.rhi = 0;
    const RSDev& R = A.rs[win.rs];
    const int32_t L0 = J.start + win.p0, L1 = J.start + win.p1, nW = win.p1 - win.p0;
    // hot streams of this read set, held in scalar registers for the window
    const int8_t* __restrict__ gq = R.quals;
    const uint8_t* __restrict__ gb = R.bases2;
    const int64_t* __restrict__ gso = R.seq_off;
    const RMeta* __restrict__ gm = R.meta;
    {
      uint4* l4 = (uint4*)lds;   // 16-byte stores: lds is 16-byte aligned
      for (int k = tid; k < lds_words / 2; k += BLOCK) l4[k] = make_uint4(0u, 0u, 0u, 0u);
      if ((lds_words & 1) && tid == 0) lds[lds_words - 1] = 0ull;
    }
    if (tid == 0) { s_ovf = 0; s_nrec = 0; }
    __syncthreads();

    LdsSink sk;
    sk.W = W; sk.bs = bs; sk.mqq = mqq; sk.ins_ = insA; sk.del_ = delA; sk.bad_ = badA; sk.clp_ = clpA;
    sk.pc = pcA; sk.is = isA; sk.rec = A.rec; sk.rec_n = A.rec_n; sk.rec_cap = A.rec_cap;
    sk.lrec = s_rec; sk.lrec_n = &s_nrec;
    sk.overflow = A.overflow; sk.st = &A.st[win.job]; sk.bcount = 0; sk.aligned_n = 0;

    // Each wave takes its own chunks of 64 reads (wave-strided over the
    // window's read range): staging, the run scan and the base loop are
    // wave-local, so the read loop has no workgroup barrier.
    RunL* wrun = s_run + wave * (2 * 64 + 1);
    // the next chunk's read facts are loaded while the current one is piled
    int64_t t0 = win.plo + 64 * wave;
    RMeta mN;
    int64_t sN = 0, eN = 0, qN = 0;
    mN.info = 0;
    if (t0 + lane < win.hi) { mN = gm[t0 + lane]; sN = gso[t0 + lane]; eN = gso[t0 + lane + 1]; }
    if (t0 < win.hi) qN = gso[t0];
    for (; t0 < win.hi; t0 += BLOCK) {
      const int64_t qbase = qN;
      const RMeta m = mN;
      const int64_t seq0 = sN, seqE = eN;
      {
        const int64_t tn = t0 + BLOCK;
        if (tn + lane < win.hi) { mN = gm[tn + lane]; sN = gso[tn + lane]; eN = gso[tn + lane + 1]; }
        if (tn < win.hi) qN = gso[tn];
      }
      // ---- stage: one read per lane.  A simple read gives at most one run;
      // a read with clips or indels is applied below one segment per lane;
      // the rest (non-ACGT bases, too many segments) are walked.
      const int64_t r = t0 + lane;
      const int32_t qsh = (int32_t)(qbase & 3);
      Run ru0;
      ru0.n = 0;
      int nrun = 0;
      bool cplx = false, segr = false;
      uint32_t rinfo = 0, rop0 = 0;   // facts of a segmented read, for the lanes that apply its segments
      int64_t rseq0 = 0;
      if (r < win.hi) {
        if (!(m.info & RI_OUTLIER) && read_in_query(m, J)) {
          frag_event(m, J, L0, L1, sk);
          if (m.fp_hi >= L0 && m.fp_lo < L1) {
            if (seq0 - qbase >= PL_TILE_SPAN) cplx = true;
            else if (m.info & RI_SIMPLE) nrun = simple_run(m, seq0, cf, L0, L1, &ru0, &sk.aligned_n, &sk.bcount);
            else if (!(m.info & RI_EXOTIC) && ((m.info & RI_SEG_MASK) >> RI_SEG_SHIFT) < RI_SEG_MAX) {
              segr = true;
              rinfo = m.info; rop0 = m.op0; rseq0 = seq0;
            } else {
              cplx = true;
            }
          }
        }
      }
      const int32_t rlen = segr ? (int32_t)(seqE - rseq0) : 0;
      // a run that only adds badPair over a range (M bases of an invalid
      // read) goes to the difference array, not to the base loop
      if (nrun > 0 && (ru0.fl & EF_BAD)) {
        if (ru0.n > 0) { atomicAdd(&bdA[ru0.p0], 1); if (ru0.p0 + ru0.n < nW) atomicAdd(&bdA[ru0.p0 + ru0.n], -1); }
        nrun = 0;
      }
      // ---- reads with clips or indels, one lane per segment: each round
      // loads and applies the segments of as many pending reads as fit the
      // 64 lanes (usually all of the chunk's in one round).  Base-loop runs
      // are staged at wrun[64 ...] and appended after the simple runs; a
      // round whose runs would not fit, or a read of more than 64 segments,
      // joins the walked reads instead.
      int32_t ncr = 0;
      for (uint64_t pend = __ballot(segr); pend;) {
        const bool inp = (pend >> lane) & 1ull;
        const int nsg = inp ? (int)((rinfo & RI_SEG_MASK) >> RI_SEG_SHIFT) : 0;
        const int sinc = wave_scan_dpp(nsg);
        const uint64_t take = __ballot(inp && sinc <= 64);
        if (!take) {                         // over 64 segments: walked with the other complex reads
          if (lane == __ffsll((unsigned long long)pend) - 1) cplx = true;
          pend &= pend - 1;
          continue;
        }
        pend &= ~take;
        int own = -1, sidx = 0;
        for (uint64_t sm = take; sm; sm &= sm - 1) {
          const int o = __ffsll((unsigned long long)sm) - 1;
          const int no = __builtin_amdgcn_readlane(nsg, o);
          const int bo = __builtin_amdgcn_readlane(sinc, o) - no;
          if (lane >= bo && lane < bo + no) { own = o; sidx = lane - bo; }
        }
        const int src = own < 0 ? lane : own;
        const uint32_t info = (uint32_t)__shfl((int)rinfo, src, 64);
        const uint32_t op0 = (uint32_t)__shfl((int)rop0, src, 64);
        const int64_t s0 = ((int64_t)__shfl((int)(rseq0 >> 32), src, 64) << 32) |
                           (uint32_t)__shfl((int)(uint32_t)rseq0, src, 64);
        const int32_t Lc = __shfl(rlen, src, 64);
        const bool has = own >= 0;
        Seg g;
        if (has) g = R.segs[op0 + sidx];
        if (ncr + __popcll(__ballot(has && seg_has_run(cf, info, Lc, L0, L1, g))) > 64) {
          if ((take >> lane) & 1ull) cplx = true;   // runs would not fit: walked instead
          continue;
        }
        Run ru;
        int got = has ? seg_apply(J, cf, t0 + own, info, s0, Lc, L0, L1, g, ru, &sk.aligned_n, &sk.bcount, sk) : 0;
        if (got && (ru.fl & EF_BAD)) {
          atomicAdd(&bdA[ru.p0], 1);
          if (ru.p0 + ru.n < nW) atomicAdd(&bdA[ru.p0 + ru.n], -1);
          got = 0;
        }
        const uint64_t gmask = __ballot(got != 0);
        if (got) wrun[64 + ncr + __popcll(gmask & lanemask_lt())] =
            RunL{ru.n, (uint32_t)ru.p0 | ru.fl, (int32_t)(ru.q - qbase) + qsh, 0};
        ncr += __popcll(gmask);
      }
      // ---- run table: simple runs (wave scan), then the staged runs.  Runs
      // are cut in chunks of PL_CHUNK bases; .off is the run's first chunk in
      // the chunk stream, .n its length in bases.
      const uint64_t v = nrun > 0 ? (1ull << 32) + (uint64_t)((ru0.n + PL_CHUNK - 1) / PL_CHUNK) : 0ull;
      const uint64_t inc = wave_scan_dpp(v);
      const uint64_t tot = lane63(inc);
      if (nrun > 0) {
        const uint64_t ex = inc - v;
        wrun[(int32_t)(ex >> 32)] =
            RunL{(int32_t)(uint32_t)ex, (uint32_t)ru0.p0 | ru0.fl, (int32_t)(ru0.q - qbase) + qsh, ru0.n};
      }
      const int32_t Es = (int32_t)(tot >> 32);
      const int32_t Cs = (int32_t)(uint32_t)tot;
      int32_t Cc = 0;
      if (ncr > 0) {
        // every lane reads its staged run before any lane writes (one wave,
        // LDS operations in order)
        RunL cr = RunL{0, 0u, 0, 0};
        if (lane < ncr) cr = wrun[64 + lane];
        const int32_t cn = lane < ncr ? (cr.off + PL_CHUNK - 1) / PL_CHUNK : 0;
        const int32_t ci = wave_scan_dpp(cn);
        Cc = lane63(ci);
        if (lane < ncr) wrun[Es + lane] = RunL{Cs + ci - cn, cr.pi, cr.q0, cr.off};
      }
      const int32_t E = Es + ncr;
      const int32_t CT = Cs + Cc;
      if (lane == 0) wrun[E] = RunL{CT, 0u, 0, 0};
      // LDS operations of one wave complete in order: the table written
      // above is what the loop below reads
      __builtin_amdgcn_wave_barrier();
      // ---- the bases of all runs: one chunk of PL_CHUNK consecutive bases of
      // one run per lane.  A chunk's quals come in as PL_CHUNK/4 + 2 dwords
      // and its 2-bit bases as PL_CHUNK/16 + 2, realigned with funnel shifts
      // (the loads of the next batch of chunks are in flight while the
      // current batch does its atomics); each base then costs a
      // byte and a 2-bit extract and two LDS atomics whose step offset is the
      // instruction's immediate.  Lanes of one run are PL_CHUNK (odd)
      // positions apart, so their 64-bit atomics spread over all LDS banks.
      if (CT > 0) {
        const int32_t mq = cf.min_qual;
        const uint32_t W8 = 8u * (uint32_t)W;
        const int64_t qb4 = qbase & ~(int64_t)3, qb16 = qbase & ~(int64_t)15;
        const gu32* tq32 = (const gu32*)(gq + qb4);
        const gu32* tb32 = (const gu32*)(gb + (qb16 >> 2));
        const uint32_t qd = (uint32_t)(qb4 - qb16);
        // run of chunk cb + lane = the last run whose first chunk is <= it
        auto run_of = [&](int32_t cb, int cur) -> int {
          const int32_t C = cb + lane;
          const int32_t lim = CT < cb + 64 ? CT : cb + 64;
          int rr = cur;
          for (int c = cur;;) {
            const int ib = c + 1 + lane;
            const int32_t bnd = ib <= E ? wrun[ib].off : INT32_MAX;
            const int nb = __popcll(__ballot(bnd < lim));
            for (int i = 0; i < nb; i++) rr += (C >= __builtin_amdgcn_readlane(bnd, i)) ? 1 : 0;
            if (nb < 64) break;
            c += 64;
          }
          return rr;
        };
        // sequence offset of chunk C of run ru (0 for a lane past the end)
        auto chunk_q = [&](const RunL& ru, int32_t C) -> uint32_t {
          return C < CT ? (uint32_t)(ru.q0 + PL_CHUNK * (C - ru.off)) : 0u;
        };
        // The loads of a batch are issued before the previous batch's
        // atomics, so they are in flight while those run.
        uint32_t qw[PL_CHUNK / 4 + 2], bw[PL_CHUNK / 16 + 2];
        auto load = [&](uint32_t q) {
#pragma unroll
          for (int k = 0; k < PL_CHUNK / 4 + 2; k++) qw[k] = tq32[(q >> 2) + k];
          const uint32_t qq = q + qd;
#pragma unroll
          for (int k = 0; k < PL_CHUNK / 16 + 2; k++) bw[k] = tb32[(qq >> 4) + k];
        };
        int cur = 0;
        int rr = run_of(0, 0);
        load(chunk_q(wrun[lane < CT ? rr : 0], lane));
        for (int32_t cb = 0; cb < CT; cb += 64) {
          const int32_t C = cb + lane;
          const bool act = C < CT;
          const RunL ru = wrun[act ? rr : cur];
          const int32_t d = PL_CHUNK * (C - ru.off);
          int32_t cnt = ru.pad - d;
          cnt = act ? (cnt < PL_CHUNK ? cnt : PL_CHUNK) : 0;
          const int32_t p = (int32_t)(ru.pi & 0xffffu) + (act ? d : 0);
          const uint32_t q = act ? (uint32_t)(ru.q0 + d) : 0u;
          const int32_t mq1 = (int32_t)((ru.pi >> 16) & 0x1ffu);
          const uint32_t qq = q + qd;
          uint32_t w[PL_CHUNK / 4 + 1];
#pragma unroll
          for (int k = 0; k < PL_CHUNK / 4 + 1; k++) w[k] = __builtin_amdgcn_alignbyte(qw[k + 1], qw[k], q & 3u);
          uint32_t x[PL_CHUNK / 16 + 1];
#pragma unroll
          for (int k = 0; k < PL_CHUNK / 16 + 1; k++) x[k] = __builtin_amdgcn_alignbit(bw[k + 1], bw[k], 2u * (qq & 15u));
          // the next batch: its runs, and its loads
          int rrn = rr;
          uint32_t qn = 0u;
          if (cb + 64 < CT) {
            const int curn = lane63(rr);
            rrn = run_of(cb + 64, curn);
            qn = chunk_q(wrun[C + 64 < CT ? rrn : curn], C + 64);
          }
          load(qn);
          unsigned long long* bsp = bs + p;
          unsigned long long* mqp = mqq + p;
          // The base's counter array is b*W entries on: a 24-bit multiply-add
          // (b < 4, 8*W < 2^24).  With minQual >= 0 every qual that passes is
          // non-negative, so qual*mq1 >= 0 and the count part is the constant
          // 1 << PL_BASE_SHIFT (pack_base without its sign extension).
          auto pile = [&](auto nonneg) {
#pragma unroll
            for (int t = 0; t < PL_CHUNK; t++) {
              const int32_t qual = (int32_t)(int8_t)(w[t >> 2] >> (8 * (t & 3)));
              const uint32_t b = (x[t >> 4] >> (2 * (t & 15))) & 3u;
              if (t < cnt && qual >= mq) {
                unsigned long long* cell = (unsigned long long*)((char*)bsp + __umul24(b, W8)) + t;
                const unsigned long long v = decltype(nonneg)::value
                                                 ? ((1ull << PL_BASE_SHIFT) | (uint32_t)(qual * mq1))
                                                 : (unsigned long long)pack_base(qual, mq1);
                atomicAdd(cell, v);
                atomicAdd(mqp + t, (unsigned long long)pack_mq_q(mq1, qual));
              }
            }
          };
          if (mq >= 0) pile(std::integral_constant<bool, true>{});
          else pile(std::integral_constant<bool, false>{});
          cur = lane63(rr);
          rr = rrn;
        }
      }
      // ---- everything else about the chunk's complex reads, one at a time
      for (uint64_t cm = __ballot(cplx); cm; cm &= cm - 1) {
        const int l = __ffsll((unsigned long long)cm) - 1;
        walk_read(R, J, cf, t0 + l, L0, L1, lane, 64, sk, false);
      }
      __builtin_amdgcn_wave_barrier();
    }
    // reads that reach the window only through their fragments: 8 bytes
    // each, +1/+size at position 0 summed per wave
    {
      // read the descriptors again here rather than hold them through the read loop
      asm volatile("" ::: "memory");
      const FragTail* __restrict__ ft = R.ftail;
      const int64_t flo = win.lo, fhi = win.plo;
      uint32_t c0 = 0, s0 = 0;
      for (int64_t r = flo + tid; r < fhi; r += BLOCK) {
        const FragTail t = ft[r];
        if (frag_tail_event(t, L0, L1, sk)) { c0++; s0 += (uint32_t)t.size; }
      }
      c0 = wave_total_dpp(c0);
      s0 = wave_total_dpp(s0);
      if (lane == 0 && c0) sk.frag(0, (int32_t)c0, (int32_t)s0);
    }
    for (int64_t k = wave; k < R.n_outlier; k += NW) {
      const int64_t r = uniform64(R.outliers[k]);
      walk_read(R, J, cf, r, L0, L1, lane, 64, sk, true);
    }
    if (tid == 0) {
      s_slot[cur ^ 1] = persistent ? next_slot : A.n_wins;
      if (persistent && next_slot < A.n_wins) s_head[cur ^ 1] = A.heads[next_slot];
    }
    __syncthreads();

    // inclusive scan of the physCov / insertSize differences over the window
    {
      const int per = (W + BLOCK - 1) / BLOCK;
      const int b0 = tid * per;
      int32_t spc = 0, sis = 0, sbd = 0;
      for (int i = 0; i < per; i++)
        if (b0 + i < W) {
          spc = j_add(spc, pcA[b0 + i]); sis = j_add(sis, isA[b0 + i]); sbd = j_add(sbd, bdA[b0 + i]);
        }
      // (Int adds wrap like Java's: the scan's 32-bit adds are the same)
      const int32_t ipc = wave_scan_dpp(spc), iis = wave_scan_dpp(sis), ibd = wave_scan_dpp(sbd);
      if (lane == 63) { ws_a[wave] = ipc; ws_b[wave] = iis; ws_c[wave] = ibd; }
      __syncthreads();
      int32_t cpc = j_sub(ipc, spc), cis = j_sub(iis, sis), cbd = j_sub(ibd, sbd);
      for (int w2 = 0; w2 < wave; w2++) {
        cpc = j_add(cpc, ws_a[w2]); cis = j_add(cis, ws_b[w2]); cbd = j_add(cbd, ws_c[w2]);
      }
      for (int i = 0; i < per; i++) {
        if (b0 + i >= W) break;
        cpc = j_add(cpc, pcA[b0 + i]);
        cis = j_add(cis, isA[b0 + i]);
        cbd = j_add(cbd, bdA[b0 + i]);
        pcA[b0 + i] = cpc;
        isA[b0 + i] = cis;
        badA[b0 + i] += (uint32_t)cbd;
      }
    }
    __syncthreads();

    // the window's evidence records: one atomic, coalesced 32-byte stores
    {
      const uint32_t nl = s_nrec < PL_WIN_RECS ? s_nrec : PL_WIN_RECS;
      if (tid == 0 && nl) s_recbase = atomicAdd(A.rec_n, nl);
      __syncthreads();
      for (uint32_t t = tid; t < nl; t += BLOCK) {
        const uint32_t k = s_recbase + t;
        if (k < A.rec_cap) A.rec[k] = s_rec[t];
        else *A.overflow = 1;
      }
    }
    const int64_t G = A.G;
    uint32_t wmask = 0;   // bit i: position tid + i*BLOCK is wide
    for (int p = tid, it = 0; p < nW; p += BLOCK, it++) {
      const uint64_t b4[4] = {bs[p], bs[W + p], bs[2 * W + p], bs[3 * W + p]};
      PosOut o;
      finalize_position(b4, mqq[p], insA[p], delA[p], badA[p], clpA[p], pcA[p], isA[p], R.ref[L0 + p - 1], cf, o);
      const int64_t g = J.out + win.p0 + p;
#pragma unroll
      for (int f = 0; f < PL_NFIELDS; f++) A.pile[(int64_t)f * G + g] = o.f[f];
      if (o.wide) { o.call.flags |= PL_CF_WIDE; wmask |= 1u << it; }
      A.calls[g] = o.call;
      if (o.overflow) s_ovf = 1;
      const int d = call_disp(o.call);
      const int64_t dep = (int64_t)(uint32_t)o.f[0] + (uint32_t)o.f[1] + (uint32_t)o.f[2] + (uint32_t)o.f[3] +
                          o.f[PL_F_DELETIONS];
      const bool is_cand = !o.defer && (d == PL_D_SNP || d == PL_D_AMB);
      if (!o.defer && d == PL_D_CONFIRMED) atomicAdd(&hist[dep > 255 ? 255 : (dep < 0 ? 0 : dep)], 1u);
      // one global atomic per wave for each list
      const uint32_t kd = wave_append(A.defer_n, o.defer);
      const uint32_t kc = wave_append(A.cand_n, is_cand);
      if (o.defer) {
        if (kd < A.defer_cap) A.defer[kd] = (uint32_t)g;
        else *A.overflow = 1;
      }
      if (is_cand) {
        if (kc < A.cand_cap) { Cand c; c.gpos = (uint32_t)g; c.depth = (int32_t)dep; c.kind = d; c.len = 0; c.rec = -1; A.cand[kc] = c; }
        else *A.overflow = 1;
      }
    }
    // wide columns (a qualSum beyond int32): the window's entries, in one
    // segment of the batch's list, with their exact sums from the LDS words
    if (__syncthreads_or(wmask != 0)) {
      const uint32_t mine = (uint32_t)__popc(wmask);
      const uint32_t inc = wave_scan_dpp(mine);
      if (lane == 63) ws_a[wave] = (int32_t)inc;
      __syncthreads();
      uint32_t pre = inc - mine, tot = 0;
      for (int w2 = 0; w2 < NW; w2++) { if (w2 < wave) pre += (uint32_t)ws_a[w2]; tot += (uint32_t)ws_a[w2]; }
      if (tid == 0) {
        s_recbase = atomicAdd(A.wide_n, tot);
        A.win_wide[2 * wi] = (int32_t)s_recbase;
        A.win_wide[2 * wi + 1] = (int32_t)tot;
      }
      __syncthreads();
      for (uint32_t m = wmask; m; m &= m - 1) {
        const int p = tid + (__ffs(m) - 1) * BLOCK;
        const uint32_t k = s_recbase + pre++;
        if (k < A.wide_cap) {
          WideCol wc;
          wc.gpos = (uint32_t)(J.out + win.p0 + p);
          wc.pad = 0;
          for (int b = 0; b < 4; b++) { uint32_t n; unpack_base(bs[b * W + p], &wc.qs[b], &n); }
          A.wide[k] = wc;
        } else {
          *A.overflow = 1;
        }
      }
    } else if (tid == 0) {
      A.win_wide[2 * wi] = 0;
      A.win_wide[2 * wi + 1] = 0;
    }
    // region baseCount and the aligned-base unit
    unsigned long long bc = wave_total_dpp((unsigned long long)sk.bcount);
    unsigned long long al = wave_total_dpp((unsigned long long)sk.aligned_n);
    if (lane == 0) { red[0][wave] = bc; red[1][wave] = al; }
    __syncthreads();
    JobStat& S = A.st[win.job];
    if (tid == 0) {
      unsigned long long t0s = 0, t1s = 0;
      for (int w2 = 0; w2 < NW; w2++) { t0s += red[0][w2]; t1s += red[1][w2]; }
      if (t0s) atomicAdd((unsigned long long*)&S.base_count, t0s);
      if (t1s) atomicAdd((unsigned long long*)&S.aligned, t1s);
      if (s_ovf) atomicOr(&S.overflow, 1);
    }
    for (int d = tid; d < 256; d += BLOCK)
      if (hist[d]) atomicAdd((uint32_t*)&S.hist[d], hist[d]);
    __syncthreads();
  }
}

template __global__ void k_window<256>(EngineArgs, int);
template __global__ void k_window<512>(EngineArgs, int);
template __global__ void k_window<1024>(EngineArgs, int);

// ------------------------------------------------------------------ records
__global__ void k_rec_keys(const IndelRec* rec, const uint32_t* rec_n, uint32_t cap, uint32_t* key, uint32_t* val) {
  const uint32_t n = *rec_n < cap ? *rec_n : cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    key[i] = rec[i].key;
    val[i] = i;
  }
}
__global__ void k_cand_keys(const Cand* c, const uint32_t* cand_n, uint32_t cap, uint32_t* key, uint32_t* val) {
  const uint32_t n = *cand_n < cap ? *cand_n : cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    key[i] = c[i].gpos;
    val[i] = i;
  }
}
__global__ void k_cand_gather(const Cand* in, const uint32_t* cand_n, uint32_t cap, const uint32_t* val, Cand* out) {
  const uint32_t n = *cand_n < cap ? *cand_n : cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = in[val[i]];
}

// ------------------------------------------------------------------ radix sort
__global__ void __launch_bounds__(256) k_rs_hist(const uint32_t* keys, const uint32_t* n_ptr, uint32_t cap, int shift,
                                                 uint32_t* hist, int32_t nblocks) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t n = *n_ptr < cap ? *n_ptr : cap;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int i = 0; i < RS_ITEMS; i++) {
    const int64_t idx = base + (int64_t)i * 256 + threadIdx.x;
    if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_rs_scatter(const uint32_t* kin, const uint32_t* vin, uint32_t* kout, uint32_t* vout,
                                                    const uint32_t* n_ptr, uint32_t cap, int shift,
                                                    const uint32_t* off, int32_t nblocks) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t wc[4][256];
  const int tid = threadIdx.x, wave = tid >> 6;
  run[tid] = off[(int64_t)tid * nblocks + blockIdx.x];
  for (int w = 0; w < 4; w++) wc[w][tid] = 0;
  __syncthreads();
  const uint32_t n = *n_ptr < cap ? *n_ptr : cap;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int round = 0; round < RS_ITEMS; round++) {
    const int64_t idx = base + (int64_t)round * 256 + tid;
    const bool valid = idx < n;
    const uint32_t k = valid ? kin[idx] : 0u;
    const uint32_t v = valid ? vin[idx] : 0u;
    const uint32_t d = (k >> shift) & 255u;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = ((d >> b) & 1u) != 0;
      const uint64_t bb = __ballot(valid && bit);
      m &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(m & lanemask_lt());
    if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(m);
    __syncthreads();
    if (valid) {
      uint32_t pre = run[d];
      for (int w = 0; w < wave; w++) pre += wc[w][d];
      kout[pre + rank] = k;
      vout[pre + rank] = v;
    }
    __syncthreads();
    run[tid] += wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
    for (int w = 0; w < 4; w++) wc[w][tid] = 0;
    __syncthreads();
  }
}

// ------------------------------------------------------------------ indel pass
namespace {
__device__ __forceinline__ int32_t job_of(const JobDev* jobs, int32_t n_jobs, int64_t g) {
  int32_t lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    int32_t m = (lo + hi + 1) >> 1;
    if (jobs[m].out <= g) lo = m; else hi = m - 1;
  }
  return lo;
}
}  // namespace

namespace {

// first index in keys[0, n) whose key is >= x, by the whole wave: 64 probes
// per round cut the range 65-fold (4 dependent loads for a million keys)
__device__ int64_t wave_lower_bound(const uint32_t* keys, int64_t n, uint32_t x, int lane) {
  int64_t lo = 0, hi = n;
  while (hi - lo > 64) {
    const int64_t m = hi - lo;
    const int64_t probe = lo + ((int64_t)(lane + 1) * m) / 65;
    const int k = __popcll(__ballot(keys[probe] < x));   // probes ascend: a prefix of lanes
    const int64_t plo = lo + ((int64_t)k * m) / 65;       // probe of lane k-1 (lo when k = 0)
    const int64_t phi = lo + ((int64_t)(k + 1) * m) / 65; // probe of lane k (hi when k = 64)
    if (k > 0) lo = plo + 1;
    hi = k < 64 ? phi : hi;
  }
  const int64_t i = lo + lane;
  return lo + __popcll(__ballot(i < hi && keys[i] < x));
}

// strict majority over one (position, kind) group of sorted records, by the
// whole wave.  Each lane runs a Boyer-Moore vote over its strided share of
// the records, comparing (length, hash) classes; the 64 partial votes are
// merged pairwise (a merge keeps a majority class if there is one), and the
// winning class's string is counted exactly.  A class with a strict
// majority whose representative string falls short (a hash collision) is
// resolved by the serial scan.  Same result as group_majority.
__device__ Majority wave_majority(const RSDev& R, const IndelRec* recs, const uint32_t* idx, int64_t a, int64_t b,
                                  int lane) {
  Majority m;
  m.n = (int32_t)(b - a);
  m.win = 0;
  m.rec = -1;
  if (b <= a) return m;
  int32_t vl = -1, vc = 0;
  uint32_t vh = 0;
  int64_t vr = -1;
  for (int64_t i = a + lane; i < b; i += 64) {
    const int64_t r = idx[i];
    const int32_t len = recs[r].len;
    const uint32_t h = recs[r].hash;
    if (vc == 0) { vl = len; vh = h; vr = r; vc = 1; }
    else if (len == vl && h == vh) vc++;
    else vc--;
  }
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t ol = __shfl_xor(vl, d, 64), oc = __shfl_xor(vc, d, 64);
    const uint32_t oh = (uint32_t)__shfl_xor((int)vh, d, 64);
    const int64_t orr = __shfl_xor((long long)vr, d, 64);
    if (ol == vl && oh == vh) vc += oc;
    else if (vc >= oc) vc -= oc;
    else { vl = ol; vh = oh; vr = orr; vc = oc - vc; }
  }
  const int64_t cr = __shfl((long long)vr, 0, 64);
  if (cr < 0) return m;
  const IndelRec c = recs[cr];
  int32_t cls = 0, same = 0;
  for (int64_t i = a + lane; i < b; i += 64) {
    const IndelRec& e = recs[idx[i]];
    if (e.len == c.len && e.hash == c.hash) {
      cls++;
      if (rec_same(R, e, c)) same++;
    }
  }
  cls = wave_sum(cls);
  same = wave_sum(same);
  if (2 * same > m.n) { m.win = same; m.rec = cr; return m; }
  if (2 * cls > m.n) return group_majority(R, recs, idx, a, b);   // collision inside the majority class
  m.win = same;                                                     // no strict majority
  return m;
}

// the region (job) holding batch position g: the last job whose first
// position is <= g, by the whole wave as in wave_lower_bound
__device__ int32_t wave_job_of(const JobDev* jobs, int32_t n_jobs, int64_t g, int lane) {
  int32_t lo = 0, hi = n_jobs;                          // first job with out > g lies in [lo, hi]
  while (hi - lo > 64) {
    const int64_t m = hi - lo;
    const int32_t probe = lo + (int32_t)(((int64_t)(lane + 1) * m) / 65);
    const int k = __popcll(__ballot(jobs[probe].out <= g));
    const int32_t plo = lo + (int32_t)(((int64_t)k * m) / 65);
    const int32_t phi = lo + (int32_t)(((int64_t)(k + 1) * m) / 65);
    if (k > 0) lo = plo + 1;
    if (k < 64) hi = phi;
  }
  const int32_t i = lo + lane;
  return lo + __popcll(__ballot(i < hi && jobs[i].out <= g)) - 1;
}

}  // namespace

// One wave per listed column: the full BaseCall with the column's evidence
// records (hetIndelCall), lane 0 writing the results.
// mode 0: deferred columns (writes call, appends candidates, counts confirmed)
// mode 1: columns changed by the deletion pass (keeps disposition bits)
__global__ void __launch_bounds__(256) k_indel_pass(EngineArgs A, const uint32_t* list, const uint32_t* list_n,
                                                     uint32_t cap, int mode) {
  const uint32_t n = *list_n < cap ? *list_n : cap;
  const uint32_t nrec = *A.rec_n < A.rec_cap ? *A.rec_n : A.rec_cap;
  const int lane = threadIdx.x & 63;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
    const uint32_t g = list[i];
    const int32_t j = wave_job_of(A.jobs, A.n_jobs, g, lane);
    const JobDev& J = A.jobs[j];
    const RSDev& R = A.rs[J.rs];
    Col c = load_col(A.pile, A.G, g);
    const uint16_t wflag = A.calls[g].flags & PL_CF_WIDE;
    if (wflag) wide_patch(c, A.wide, A.wide_cap, A.win_wide, J, A.W, g);
    const uint8_t refb = R.ref[J.start + (int64_t)(g - J.out) - 1];
    const BaseFacts bf = base_facts(c);
    int32_t pct = 0;
    const int need = indel_kind_needed(c, A.cf, &pct);
    int indel = 0;
    bool homo = true;
    int32_t ilen = 0;
    int64_t wrec = -1;
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
      const Majority mj = wave_majority(R, A.rec, A.rval, a, b, lane);
      if (mj.rec >= 0) {
        const IndelRec& e = A.rec[mj.rec];
        const int d = het_indel_decide(mj.n, mj.win, e.len, (e.flags & 2u) != 0, pct, A.cf);
        if (d) { indel = need; homo = d == 2; ilen = e.len; wrec = mj.rec; }
      }
    }
    if (lane != 0) continue;
    pl_call cl = make_call(bf, indel, homo, refb, A.cf);
    cl.flags |= wflag;
    if (mode == 0) {
      A.calls[g] = cl;
      const int d = call_disp(cl);
      const int64_t dep = c.depth();
      if (d == PL_D_CONFIRMED) atomicAdd((uint32_t*)&A.st[j].hist[dep > 255 ? 255 : (dep < 0 ? 0 : dep)], 1u);
      else if (d != PL_D_NONE) {
        uint32_t k = atomicAdd(A.cand_n, 1u);
        if (k < A.cand_cap) { Cand cd; cd.gpos = g; cd.depth = (int32_t)dep; cd.kind = d; cd.len = ilen; cd.rec = wrec; A.cand[k] = cd; }
        else *A.overflow = 1;
      }
    } else {
      pl_call& o = A.calls[g];
      o.base = cl.base; o.alt = cl.alt; o.q = cl.q;
      o.flags = (uint16_t)((cl.flags & ~PL_CF_DISP_MASK) | (o.flags & PL_CF_DISP_MASK) | PL_CF_DELETED);
    }
  }
}

// ------------------------------------------------------------------ Pass 1
// GenomeRegion.postProcess Pass 1 (GenomeRegion.scala:221-272) for one
// region per wave.  Only deletion calls carry state along the scan (a
// called deletion suppresses the candidates under it), so: the region's
// deletion candidates are scanned in order to find the active ones (few);
// every candidate is then cleared at once if too shallow or under an active
// deletion; each active deletion's positions are then mutated by the lanes.
// Same result as finalize_job, which a region with more than PL_FIN_DELS
// active deletions runs serially.
__global__ void __launch_bounds__(64) k_finalize(EngineArgs A, const Cand* sorted, const uint32_t* cand_n, uint32_t cap,
                                                 Cand* cand) {
  __shared__ int64_t s_dg[PL_FIN_DELS];
  __shared__ int32_t s_dl[PL_FIN_DELS];
  const int32_t j = blockIdx.x;
  if (j >= A.n_jobs) return;
  const int lane = threadIdx.x;
  const uint32_t n = *cand_n < cap ? *cand_n : cap;
  const JobDev& J = A.jobs[j];
  JobStat& st = A.st[j];
  // candidates of this region: gpos in [out, out + size)
  int64_t a, b;
  {
    int64_t lo = 0, hi = n;
    while (hi - lo > 64) {
      const int64_t m = hi - lo, pr = lo + ((int64_t)(lane + 1) * m) / 65;
      const int k = __popcll(__ballot((int64_t)sorted[pr].gpos < J.out));
      const int64_t plo = lo + ((int64_t)k * m) / 65, phi = lo + ((int64_t)(k + 1) * m) / 65;
      if (k > 0) lo = plo + 1;
      if (k < 64) hi = phi;
    }
    a = lo + __popcll(__ballot(lo + lane < hi && (int64_t)sorted[lo + lane].gpos < J.out));
    lo = a; hi = n;
    while (hi - lo > 64) {
      const int64_t m = hi - lo, pr = lo + ((int64_t)(lane + 1) * m) / 65;
      const int k = __popcll(__ballot((int64_t)sorted[pr].gpos < J.out + J.size));
      const int64_t plo = lo + ((int64_t)k * m) / 65, phi = lo + ((int64_t)(k + 1) * m) / 65;
      if (k > 0) lo = plo + 1;
      if (k < 64) hi = phi;
    }
    b = lo + __popcll(__ballot(lo + lane < hi && (int64_t)sorted[lo + lane].gpos < J.out + J.size));
  }
  if (lane == 0) { st.cand_lo = (int32_t)a; st.cand_hi = (int32_t)b; }
  const int32_t md = min_depth_of(A.cf, st.base_count, J.size);
  const bool none = st.read_count == 0;
  // active deletions, in order
  int nd = 0;
  int64_t delEnd = -1;
  bool serial = false;
  for (int64_t c0 = a; c0 < b && !serial; c0 += 64) {
    const int64_t k = c0 + lane;
    const bool isd = k < b && !none && cand[k].kind == PL_D_DEL && cand[k].depth >= md;
    for (uint64_t m = __ballot(isd); m; m &= m - 1) {
      const int l = __ffsll((unsigned long long)m) - 1;
      const int64_t g = (int64_t)cand[c0 + l].gpos;
      const int32_t len = cand[c0 + l].len;
      if (g <= delEnd) continue;
      if (nd == PL_FIN_DELS) { serial = true; break; }
      if (lane == 0) { s_dg[nd] = g; s_dl[nd] = len; }
      nd++;
      delEnd = g + len - 1;
    }
  }
  __syncthreads();
  if (serial) {
    if (lane == 0)
      finalize_job(A.cf, J, st, A.pile, A.G, A.calls, cand, a, b, A.mutated, A.mutated_n, A.mutated_cap, A.overflow);
    return;
  }
  // every candidate at once: shallow, or under an active deletion (not its own)
  for (int64_t c0 = a; c0 < b; c0 += 64) {
    const int64_t k = c0 + lane;
    if (k >= b) continue;
    const int64_t g = (int64_t)cand[k].gpos;
    int lo = 0, hi = nd;                                  // active deletions starting before g
    while (lo < hi) { const int mid = (lo + hi) >> 1; if (s_dg[mid] < g) lo = mid + 1; else hi = mid; }
    const bool shadowed = lo > 0 && g <= s_dg[lo - 1] + s_dl[lo - 1] - 1;
    if (none || shadowed || cand[k].depth < md) cand[k].kind = 0;
  }
  // the active deletions: deleted positions take the deletion count (Pass 1
  // mutation, GenomeRegion.scala:262-264), confirmed positions under them leave the count
  int64_t dropped = 0;
  for (int i = 0; i < nd; i++) {
    const int64_t g = s_dg[i];
    const int32_t len = s_dl[i];
    const int32_t dcount = A.pile[(int64_t)PL_F_DELETIONS * A.G + g];
    for (int32_t q = 1 + lane; q < len; q += 64) {
      const int64_t gj = g + q;
      int64_t dep = 0;
      for (int bb = 0; bb < 4; bb++) dep += (int64_t)(uint32_t)A.pile[(int64_t)(PL_F_COUNT_A + bb) * A.G + gj];
      dep += A.pile[(int64_t)PL_F_DELETIONS * A.G + gj];
      if (call_disp(A.calls[gj]) == PL_D_CONFIRMED && dep >= md) dropped++;
      A.pile[(int64_t)PL_F_DELETIONS * A.G + gj] = j_add(A.pile[(int64_t)PL_F_DELETIONS * A.G + gj], dcount);
      A.calls[gj].flags |= PL_CF_DELETED;
      const uint32_t slot = atomicAdd(A.mutated_n, 1u);
      if (slot < A.mutated_cap) A.mutated[slot] = (uint32_t)gj;
      else *A.overflow = 1;
    }
  }
  dropped = wave_sum(dropped);
  if (lane == 0) {
    st.min_depth = md;
    int64_t nconf = 0;   // minDepth above the histogram: counted by k_confirm_exact
    if (md <= PL_CONF_HIST_MAX) for (int d = md; d < 256; d++) nconf += st.hist[d];
    else if (st.read_count != 0) atomicOr(A.deep, 1u);
    if (nconf > 0) nconf = nconf - dropped > 0 ? nconf - dropped : 0;   // the serial scan stops decrementing at 0
    st.n_confirmed = (int32_t)nconf;
  }
}

// Exact confirmed count of the regions whose minDepth lies above the depth
// histogram (a --mindepth of 256 or more, or a region averaging more than
// about 2,555x): every position of their windows is tested after the
// deletion pass.  Other regions' windows are skipped after one load, so the
// launch costs a few microseconds when no region needs it.
__global__ void __launch_bounds__(256) k_confirm_exact(EngineArgs A) {
  __shared__ uint32_t s_n;
  if (*(volatile uint32_t*)A.deep == 0) return;   // no region needs the exact count
  for (int32_t w = blockIdx.x; w < A.n_wins; w += gridDim.x) {
    const WinDev win = A.wins[w];
    JobStat& S = A.st[win.job];
    const int32_t md = S.min_depth;
    if (md <= PL_CONF_HIST_MAX || S.read_count == 0) continue;
    const JobDev& J = A.jobs[win.job];
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    uint32_t n = 0;
    for (int p = win.p0 + threadIdx.x; p < win.p1; p += blockDim.x) {
      const int64_t g = J.out + p;
      int64_t dep = 0;
      for (int b = 0; b < 4; b++) dep += (uint32_t)A.pile[(int64_t)(PL_F_COUNT_A + b) * A.G + g];
      dep += A.pile[(int64_t)PL_F_DELETIONS * A.G + g];
      n += confirmed_final(A.calls[g], dep, md, S.read_count) ? 1u : 0u;
    }
    n = wave_sum(n);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(&s_n, n);
    __syncthreads();
    if (threadIdx.x == 0 && s_n) atomicAdd(&S.n_confirmed, (int32_t)s_n);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ export
// final dispositions of one window's positions (GenomeRegion.scala:255)
__global__ void k_final_flags(EngineArgs A) {
  const WinDev win = A.wins[blockIdx.x];
  const JobDev& J = A.jobs[win.job];
  const JobStat& S = A.st[win.job];
  for (int p = win.p0 + threadIdx.x; p < win.p1; p += blockDim.x) {
    const int64_t g = J.out + p;
    int64_t dep = 0;
    for (int b = 0; b < 4; b++) dep += (uint32_t)A.pile[(int64_t)(PL_F_COUNT_A + b) * A.G + g];
    dep += A.pile[(int64_t)PL_F_DELETIONS * A.G + g];
    uint16_t f = A.calls[g].flags;
    const int cand = (f & PL_CF_DISP_MASK) >> PL_CF_DISP_SHIFT;
    int disp = cand;
    if (S.read_count == 0 || (f & PL_CF_DELETED) || dep < S.min_depth) disp = PL_D_NONE;
    A.calls[g].flags = (uint16_t)((f & ~PL_CF_DISP_MASK) | (disp << PL_CF_DISP_SHIFT));
  }
}

// insertion strings of INS changes into a byte pool
// the calls of the candidate columns, in candidate order (one copy to the host)
__global__ void k_cand_calls(const Cand* c, uint32_t n, const pl_call* calls, pl_call* out) {
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = calls[c[k].gpos];
}

__global__ void k_ins_strings(EngineArgs A, const Cand* sorted, const int64_t* idx, const int64_t* off, int32_t n, uint8_t* pool) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Cand& c = sorted[idx[i]];
  const IndelRec& e = A.rec[c.rec];
  const int32_t j = job_of(A.jobs, A.n_jobs, c.gpos);
  const RSDev& R = A.rs[A.jobs[j].rs];
  for (int32_t q = 0; q < e.len; q++) pool[off[i] + q] = rec_char(R, e, q);
}

#include "pl_vcf_kernels.hip"
#include "pl_track_kernels.hip"
