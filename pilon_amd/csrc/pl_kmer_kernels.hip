// pl_kmer_kernels.hip — the K-mer pileups of Pilon's local assembler
// (Assembler.addReads / addToPileups / buildGraph, Assembler.scala:46-85,
// 131-153) on the device.  Included at the end of pl_kernels.hip.
//
//   k_kmer_emit   one workgroup per read, one thread per K-mer occurrence
//                 (strand, offset): the K-mer as 4-bit BAM codes packed in
//                 six 32-bit words (first base most significant), their
//                 hash, and the base after it with its quality and the
//                 read's MAPQ;
//   (an LSD radix sort of the occurrences by a 64-bit hash of the six
//   words, two 32-bit keys; the six words themselves only if two distinct
//   K-mers share a hash)
//   k_kmer_heads  the first occurrence of each distinct K-mer (then an
//                 exclusive scan numbers the K-mers, k_kmer_starts lists
//                 where each begins), and a flag when equal hashes hold
//                 different K-mers;
//   k_kmer_call   one thread per K-mer: PileUp.add of its occurrences and,
//                 at depth >= minDepth, BaseCall: kGraph's base and, when
//                 not homozygous, altGraph's.

#define PL_KMER_WORDS 6          // K <= 48: eight 4-bit codes per word
#define PL_KMER_STRIDE 8         // words per occurrence in kw: one 32-byte record (unused words zero)

struct KmerArgs {
  int64_t n_reads;
  const int64_t* seq_off;        // [n_reads + 1]
  const uint8_t* bases2;
  const int8_t* quals;
  const uint16_t* flag;
  const uint8_t* mapq;
  const int64_t* exo_of;         // [n_reads] first nibble of a non-ACGT read, -1 otherwise
  const uint8_t* nib;
  const int64_t* occ_off;        // [n_reads + 1] first occurrence of each read
  uint32_t n_occ;
  int32_t K, words;
  int32_t default_qual, min_qual, min_depth;
  uint32_t* kw;                  // [n_occ][PL_KMER_STRIDE] each occurrence's words together
  uint32_t* hw;                  // [2][n_occ] the K-mer's hash, low word then high word
  uint32_t* val;                 // [n_occ] code | qual << 8 | mapq << 16
  uint32_t* collide;             // set when sorted neighbours share a hash but not their words
  int32_t by_hash;               // the order is by hash (heads compare hashes too)
  const uint32_t* perm;          // [n_occ] occurrences in K-mer order
  uint32_t* head;                // [n_occ]
  const uint32_t* seg;           // [n_occ] heads before each sorted occurrence (exclusive scan)
  uint32_t n_seg;
  uint32_t* starts;              // [n_seg + 1] first sorted occurrence of each K-mer
  uint32_t* out_n;
  uint32_t* out_kw;              // [PL_KMER_WORDS][n_seg]
  uint8_t* out_next;
  uint8_t* out_alt;
};

namespace {

// 4-bit BAM code of base j of read r (getReadString)
__device__ __forceinline__ uint32_t kmer_code(const KmerArgs& A, int64_t r, int64_t j) {
  const int64_t x = A.exo_of[r];
  if (x >= 0) {
    const int64_t k = x + j;
    return (A.nib[k >> 1] >> ((k & 1) ? 0 : 4)) & 15u;
  }
  const int64_t k = A.seq_off[r] + j;
  return 1u << ((A.bases2[k >> 2] >> (2 * (k & 3))) & 3u);
}

// a 64-bit mix of the K-mer's words (a grouping key: equal words give
// equal hashes; k_kmer_heads checks that equal hashes hold equal words)
__device__ __forceinline__ uint64_t kmer_hash(const uint32_t* w, int words) {
  uint64_t h = 0x9e3779b97f4a7c15ull * (uint64_t)(words + 1);
  for (int k = 0; k < words; k++) {
    h = (h ^ w[k]) * 0xff51afd7ed558ccdull;
    h ^= h >> 33;
  }
  h *= 0xc4ceb9fe1a85ec53ull;
  return h ^ (h >> 29);
}

// Bases.complement (Bases.scala:47-53): A<->T, C<->G, any other code kept
__device__ __forceinline__ uint32_t kmer_comp(uint32_t c) {
  return c == 1u ? 8u : c == 8u ? 1u : c == 2u ? 4u : c == 4u ? 2u : c;
}

__device__ __forceinline__ int32_t kmer_qual(const KmerArgs& A, int64_t r, int64_t j) {
  return (A.flag[r] & PL_RF_NO_QUAL) ? A.default_qual : (int32_t)A.quals[A.seq_off[r] + j];
}

}  // namespace

// the occurrence's outputs: its K-mer words, their hash, the next base
__device__ __forceinline__ void kmer_store(const KmerArgs& A, int64_t t, const uint32_t* w, uint32_t next,
                                           int32_t qual, uint32_t mapq) {
  uint4* rec = (uint4*)(A.kw + (int64_t)t * PL_KMER_STRIDE);
  static_assert(PL_KMER_WORDS <= PL_KMER_STRIDE && PL_KMER_STRIDE == 8, "two 16-byte halves per occurrence");
  uint32_t v[PL_KMER_STRIDE];
#pragma unroll
  for (int k = 0; k < PL_KMER_STRIDE; k++) v[k] = (k < PL_KMER_WORDS && k < A.words) ? w[k] : 0u;
  rec[0] = make_uint4(v[0], v[1], v[2], v[3]);
  rec[1] = make_uint4(v[4], v[5], v[6], v[7]);
  const uint64_t h = kmer_hash(w, A.words);
  A.hw[t] = (uint32_t)h;
  A.hw[(int64_t)A.n_occ + t] = (uint32_t)(h >> 32);
  A.val[t] = next | ((uint32_t)(uint8_t)(int8_t)qual << 8) | (mapq << 16);
}

#define PL_KMER_LDS_BASES 8192   // reads up to this long are staged in LDS

// One workgroup per read: the read's codes, forward and reverse-complemented,
// packed eight to a 32-bit word (first code most significant) in LDS; each
// occurrence's K-mer is then six funnel shifts of that stream.  Occurrence
// 2*o + rc of the read is the K-mer at offset o of the forward (rc = 0) or
// reverse-complement (rc = 1) sequence (Assembler.addRead, Assembler.scala:65-85).
__global__ void __launch_bounds__(256) k_kmer_emit(KmerArgs A) {
  __shared__ uint32_t fw[PL_KMER_LDS_BASES / 8 + 8], rw[PL_KMER_LDS_BASES / 8 + 8];
  const int64_t r = blockIdx.x;
  const int64_t t0 = A.occ_off[r], t1 = A.occ_off[r + 1];
  if (t1 <= t0) return;
  const int64_t L = A.seq_off[r + 1] - A.seq_off[r];
  const uint32_t mapq = A.mapq[r];
  if (L > PL_KMER_LDS_BASES) {
    // a long read: each occurrence reads its K + 1 codes itself
    for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
      const int64_t local = t - t0;
      const bool rc = local & 1;
      const int64_t o = local >> 1;
      uint32_t w[PL_KMER_WORDS] = {0, 0, 0, 0, 0, 0};
      for (int j = 0; j < A.K; j++) {
        const int64_t p = o + j;
        const uint32_t c = rc ? kmer_comp(kmer_code(A, r, L - 1 - p)) : kmer_code(A, r, p);
        w[j >> 3] |= c << (28 - 4 * (j & 7));
      }
      const int64_t pn = o + A.K;
      const uint32_t next = rc ? kmer_comp(kmer_code(A, r, L - 1 - pn)) : kmer_code(A, r, pn);
      const int32_t qual = rc ? kmer_qual(A, r, L - 1 - pn) : kmer_qual(A, r, pn);
      kmer_store(A, t, w, next, qual, mapq);
    }
    return;
  }
  const int nw = (int)((L + 7) / 8);
  for (int k = threadIdx.x; k < nw + 8; k += blockDim.x) {
    uint32_t f = 0, b = 0;
    for (int j = 0; j < 8; j++) {
      const int64_t p = 8 * (int64_t)k + j;
      if (p < L) {
        f |= kmer_code(A, r, p) << (28 - 4 * j);
        b |= kmer_comp(kmer_code(A, r, L - 1 - p)) << (28 - 4 * j);
      }
    }
    fw[k] = f;
    rw[k] = b;
  }
  __syncthreads();
  const int K = A.K;
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const int64_t local = t - t0;
    const bool rc = local & 1;
    const int o = (int)(local >> 1);
    const uint32_t* sw = rc ? rw : fw;
    const int a = o >> 3, sh = 4 * (o & 7);
    uint32_t w[PL_KMER_WORDS];
#pragma unroll
    for (int k = 0; k < PL_KMER_WORDS; k++) {
      const uint32_t hi = sw[a + k], lo = sw[a + k + 1];
      w[k] = sh ? (hi << sh) | (lo >> (32 - sh)) : hi;
    }
    // codes past the K-th are zero
#pragma unroll
    for (int k = 0; k < PL_KMER_WORDS; k++) {
      const int keep = K - 8 * k;
      w[k] = keep >= 8 ? w[k] : keep <= 0 ? 0u : (w[k] & ~(0xffffffffu >> (4 * keep)));
    }
    const int pn = o + K;
    const uint32_t next = (sw[pn >> 3] >> (28 - 4 * (pn & 7))) & 15u;
    const int32_t qual = rc ? kmer_qual(A, r, L - 1 - pn) : kmer_qual(A, r, pn);
    kmer_store(A, t, w, next, qual, mapq);
  }
}

// keys[i] = the key word of the i-th occurrence in the current order: src
// holds occurrence o's word at src[o * stride]
__global__ void __launch_bounds__(256) k_kmer_gather(const uint32_t* src, uint32_t stride, const uint32_t* perm,
                                                     uint32_t n, uint32_t* keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = src[(int64_t)perm[i] * stride];
}

__global__ void __launch_bounds__(256) k_kmer_heads(KmerArgs A) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool in = i < A.n_occ;
  // each occurrence's words are gathered once: the previous sorted
  // occurrence's come from the lane before (lane 0 gathers them itself)
  const bool first = lane == 0 && in && i > 0;
  const uint32_t a = in ? A.perm[i] : 0u;
  const uint32_t b = first ? A.perm[i - 1] : 0u;
  const uint4* ra = (const uint4*)(A.kw + (int64_t)a * PL_KMER_STRIDE);
  const uint4* rb = (const uint4*)(A.kw + (int64_t)b * PL_KMER_STRIDE);
  uint4 x0 = make_uint4(0u, 0u, 0u, 0u), x1 = x0, y0 = x0, y1 = x0;
  if (in) { x0 = ra[0]; x1 = ra[1]; }
  if (first) { y0 = rb[0]; y1 = rb[1]; }
  const uint32_t xs[PL_KMER_STRIDE] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
  const uint32_t ys[PL_KMER_STRIDE] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
  uint32_t h = 0;
#pragma unroll
  for (int k = 0; k < PL_KMER_WORDS; k++) {
    uint32_t y = (uint32_t)__shfl_up((int)xs[k], 1, 64);
    if (first) y = ys[k];
    h |= (k < A.words) && xs[k] != y;
  }
  if (!in) return;
  if (i == 0) {
    h = 1;
  } else if (A.by_hash && h) {
    // sorted by hash: equal hashes with different words may interleave
    const uint32_t bp = A.perm[i - 1];
    if (A.hw[a] == A.hw[bp] && A.hw[(int64_t)A.n_occ + a] == A.hw[(int64_t)A.n_occ + bp]) atomicOr(A.collide, 1u);
  }
  A.head[i] = h;
}

// the first sorted occurrence of each K-mer (starts[n_seg] = n_occ is set by the host)
__global__ void __launch_bounds__(256) k_kmer_starts(KmerArgs A) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A.n_occ && A.head[i]) A.starts[A.seg[i]] = i;
}

// one thread per K-mer: PileUp.add (PileUp.scala:75-84) of its occurrences,
// contiguous after the sort, then buildGraph's call (Assembler.scala:131-150)
__global__ void __launch_bounds__(256) k_kmer_call(KmerArgs A) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  bool keep = false;
  Col c;
  uint32_t i0 = 0;
  if (s < A.n_seg) {
    for (int b = 0; b < 4; b++) { c.bc[b] = 0; c.qs[b] = 0; }
    c.mqSum = c.qSum = c.deletions = c.insertions = c.insQual = c.delQual = 0;
    i0 = A.starts[s];
    const uint32_t i1 = A.starts[s + 1];
    // the occurrences' values come through the permutation: eight
    // permutation loads, then eight value loads, in flight together
    for (uint32_t i = i0; i < i1; i += 8) {
      uint32_t pb[8], vb[8];
#pragma unroll
      for (int u = 0; u < 8; u++) pb[u] = i + u < i1 ? A.perm[i + u] : 0u;
#pragma unroll
      for (int u = 0; u < 8; u++) vb[u] = i + u < i1 ? A.val[pb[u]] : 0u;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (i + u >= i1) break;
        const uint32_t v = vb[u];
        const uint32_t code = v & 15u;
        const int bi = code == 1u ? 0 : code == 2u ? 1 : code == 4u ? 2 : code == 8u ? 3 : -1;
        const int32_t qual = (int32_t)(int8_t)(uint8_t)(v >> 8);
        const int32_t mq1 = (int32_t)(v >> 16) + 1;
        if (bi >= 0 && qual >= A.min_qual) {
          // each sum by a constant index, so the column stays in registers
          const int64_t qm = (int64_t)j_mul(qual, mq1);
#pragma unroll
          for (int b = 0; b < 4; b++) {
            c.bc[b] += bi == b ? 1 : 0;
            c.qs[b] += bi == b ? qm : 0;
          }
          c.mqSum = j_add(c.mqSum, mq1);
          c.qSum = j_add(c.qSum, qual);
        }
      }
    }
    keep = c.depth() >= A.min_depth;
  }
  const uint32_t k = wave_append(A.out_n, keep);
  if (!keep) return;
  const BaseFacts f = base_facts(c);
  const uint32_t p = A.perm[i0];
  for (int w = 0; w < A.words; w++) A.out_kw[(int64_t)w * A.n_seg + k] = A.kw[(int64_t)p * PL_KMER_STRIDE + w];
  A.out_next[k] = (uint8_t)"ACGTN"[f.base];
  A.out_alt[k] = f.homo ? (uint8_t)0 : (uint8_t)"ACGT"[f.alt];
}
