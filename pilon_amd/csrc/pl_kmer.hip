// pl_kmer.hip — the host side of the assembler's K-mer graph (included at
// the end of pl_engine.hip): the reads go to the device, the occurrences
// are emitted and grouped by K-mer (an LSD radix sort by a 64-bit hash of
// the K-mer; should two distinct K-mers share a hash, six stable 32-bit
// sorts by the K-mer's own words, last word first), summed and called per
// K-mer by the k_kmer_* kernels; the kept K-mers come back in byte order.

namespace {

pl_status kmer_graph(pl_ctx* c, const pl_read_soa* R, int32_t K, int32_t min_depth, char** kmers, uint8_t** next,
                     uint8_t** alt, int64_t* n_out, double* device_ms) {
  *kmers = nullptr; *next = nullptr; *alt = nullptr; *n_out = 0;
  if (K < 1 || K > 8 * PL_KMER_WORDS) return fail(PL_E_INVAL, "K = %d outside [1, %d]", K, 8 * PL_KMER_WORDS);
  const int64_t n = R->n_reads;
  if (n < 0 || (n > 0 && (!R->seq_off || !R->bases2 || !R->quals || !R->flag || !R->mapq)))
    return fail(PL_E_INVAL, "incomplete read batch");
  // occurrences per read (both strands of every offset 0 .. L-K-1), the
  // nibble offset of each non-ACGT read, and the stream sizes
  std::vector<int64_t> occ((size_t)n + 1, 0), exo((size_t)n, -1);
  int64_t nreads = 0, nib_bytes = 1;
  for (int64_t i = 0; i < n; i++) {
    const int64_t L = R->seq_off[i + 1] - R->seq_off[i];
    occ[(size_t)i + 1] = occ[(size_t)i] + (L > K ? 2 * (L - K) : 0);
    if (L > K) nreads++;
  }
  for (int64_t k = 0; k < R->n_exotic; k++) {
    const int64_t r = R->exo_read[k];
    if (r < 0 || r >= n) return fail(PL_E_INVAL, "exotic read index %lld outside the batch", (long long)r);
    exo[(size_t)r] = R->exo_off[k];
    const int64_t L = R->seq_off[r + 1] - R->seq_off[r];
    nib_bytes = std::max<int64_t>(nib_bytes, (R->exo_off[k] + L + 1) / 2);
  }
  if (nreads >= 100000)
    return fail(PL_E_INVAL, "%lld reads longer than K: Assembler.addRead would prune its map at 100,000",
                (long long)nreads);
  const int64_t N = occ[(size_t)n];
  if (N >= ((int64_t)1 << 31)) return fail(PL_E_INVAL, "too many K-mer occurrences (%lld)", (long long)N);
  if (N == 0) return PL_OK;
  const int64_t S = R->seq_off[n];
  HIPCHK(hipSetDevice(c->device));
  Lane* L = &c->main;
  hipStream_t st = L->stream;
  VcfBufs B;
  KmerArgs A;
  std::memset(&A, 0, sizeof A);
  A.n_reads = n;
  A.n_occ = (uint32_t)N;
  A.K = K;
  A.words = (K + 7) / 8;
  A.default_qual = c->pc.default_qual;
  A.min_qual = c->pc.min_qual;
  A.min_depth = min_depth;
  int64_t *dso, *dexo, *docc;
  uint8_t *db2, *dmq, *dnib;
  int8_t* dq;
  uint16_t* dfl;
  pl_status s;
  if ((s = B.get(&dso, 8 * ((size_t)n + 1))) || (s = B.get(&db2, (size_t)(S + 3) / 4 + 8)) ||
      (s = B.get(&dq, (size_t)S + 8)) || (s = B.get(&dfl, 2 * (size_t)n)) || (s = B.get(&dmq, (size_t)n)) ||
      (s = B.get(&dexo, 8 * (size_t)n)) || (s = B.get(&dnib, (size_t)nib_bytes)) || (s = B.get(&docc, 8 * ((size_t)n + 1))))
    return s;
  HIPCHK(hipMemcpyAsync(dso, R->seq_off, 8 * ((size_t)n + 1), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(db2, R->bases2, (size_t)(S + 3) / 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dq, R->quals, (size_t)S, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dfl, R->flag, 2 * (size_t)n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dmq, R->mapq, (size_t)n, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dexo, exo.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st));
  if (R->n_exotic) HIPCHK(hipMemcpyAsync(dnib, R->nib, (size_t)nib_bytes, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(docc, occ.data(), 8 * ((size_t)n + 1), hipMemcpyHostToDevice, st));
  A.seq_off = dso; A.bases2 = db2; A.quals = dq; A.flag = dfl; A.mapq = dmq; A.exo_of = dexo; A.nib = dnib;
  A.occ_off = docc;
  const uint32_t nocc = (uint32_t)N;
  uint32_t *perm, *keys, *tk, *tv, *ncount, *head, *seg;
  if ((s = B.get(&A.kw, 4 * (size_t)PL_KMER_STRIDE * (size_t)N)) || (s = B.get(&A.val, 4 * (size_t)N)) ||
      (s = B.get(&A.hw, 8 * (size_t)N)) || (s = B.get(&A.collide, 4)) ||
      (s = B.get(&perm, 4 * (size_t)N)) || (s = B.get(&keys, 4 * (size_t)N)) || (s = B.get(&tk, 4 * (size_t)N)) ||
      (s = B.get(&tv, 4 * (size_t)N)) || (s = B.get(&ncount, 4)) || (s = B.get(&head, 4 * (size_t)N)) ||
      (s = B.get(&seg, 4 * ((size_t)N + 1))))                // the scan also writes the total at [N]
    return s;
  EventPair ev;
  HIPCHK(ev.create());
  HIPCHK(hipEventRecord(ev.e0, st));
  const unsigned gb = (unsigned)((N + 255) / 256);
  k_kmer_emit<<<dim3((unsigned)n), dim3(256), 0, st>>>(A);
  HIPCHK(hipGetLastError());
  // sort the occurrences (perm) by the 32-bit keys src[0..nw) of each, the
  // least significant first: each stable sort keeps the order of the later keys
  auto sort_by = [&](const uint32_t* const* src, uint32_t stride, int nw) -> pl_status {
    k_iota<<<dim3(std::min<unsigned>(gb, 4096)), dim3(256), 0, st>>>(perm, (int32_t)N, ncount);
    for (int w = 0; w < nw; w++) {
      k_kmer_gather<<<dim3(gb), dim3(256), 0, st>>>(src[w], stride, perm, nocc, keys);
      uint32_t *rk, *rv;
      if (pl_status e = radix_sort(L, keys, perm, tk, tv, ncount, nocc, 32, &rk, &rv)) return e;
      if (rv != perm) HIPCHK(hipMemcpyAsync(perm, rv, 4 * (size_t)N, hipMemcpyDeviceToDevice, st));
    }
    return PL_OK;
  };
  A.perm = perm;
  A.head = head;
  const uint32_t* by_hash[2] = {A.hw, A.hw + (size_t)N};
  if ((s = sort_by(by_hash, 1, 2))) return s;
  A.by_hash = 1;
  HIPCHK(hipMemsetAsync(A.collide, 0, 4, st));
  k_kmer_heads<<<dim3(gb), dim3(256), 0, st>>>(A);
  uint32_t collide = 0;
  HIPCHK(hipMemcpyAsync(&collide, A.collide, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  // PILON_HIP_KMER_WORD_SORT in the environment takes this path without a
  // collision (tests compare it with the hash order's graph)
  if (collide || std::getenv("PILON_HIP_KMER_WORD_SORT")) {
    // two K-mers share a hash: order by the words themselves instead
    const uint32_t* by_words[PL_KMER_WORDS];
    for (int w = 0; w < A.words; w++) by_words[w] = A.kw + (size_t)(A.words - 1 - w);
    if ((s = sort_by(by_words, PL_KMER_STRIDE, A.words))) return s;
    A.by_hash = 0;
    k_kmer_heads<<<dim3(gb), dim3(256), 0, st>>>(A);
  }
  if ((s = scan_u32(L, head, seg, N))) return s;
  uint32_t ns = 0;                                          // distinct K-mers
  HIPCHK(hipMemcpyAsync(&ns, seg + N, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  A.seg = seg;
  A.n_seg = ns;
  if ((s = B.get(&A.starts, 4 * ((size_t)ns + 1))) || (s = B.get(&A.out_n, 4)) ||
      (s = B.get(&A.out_kw, 4 * (size_t)A.words * (size_t)ns)) || (s = B.get(&A.out_next, (size_t)ns)) ||
      (s = B.get(&A.out_alt, (size_t)ns)))
    return s;
  HIPCHK(hipMemsetAsync(A.out_n, 0, 4, st));
  HIPCHK(hipMemcpyAsync(A.starts + ns, &nocc, 4, hipMemcpyHostToDevice, st));
  k_kmer_starts<<<dim3(gb), dim3(256), 0, st>>>(A);
  k_kmer_call<<<dim3((ns + 255) / 256), dim3(256), 0, st>>>(A);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev.e1, st));
  uint32_t nk = 0;
  HIPCHK(hipMemcpyAsync(&nk, A.out_n, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<uint32_t> kw((size_t)A.words * ns);
  std::vector<uint8_t> nx(nk), al(nk);
  HIPCHK(hipMemcpyAsync(kw.data(), A.out_kw, 4 * (size_t)A.words * ns, hipMemcpyDeviceToHost, st));
  if (nk) {
    HIPCHK(hipMemcpyAsync(nx.data(), A.out_next, nk, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(al.data(), A.out_alt, nk, hipMemcpyDeviceToHost, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, ev.e0, ev.e1));
  // the K-mers as text, in byte order
  static const char NIB[] = "=ACMGRSVTWYHKDBN";
  std::string text((size_t)K * nk, ' ');
  for (uint32_t k = 0; k < nk; k++)
    for (int j = 0; j < K; j++)
      text[(size_t)K * k + j] = NIB[(kw[(size_t)(j >> 3) * ns + k] >> (28 - 4 * (j & 7))) & 15u];
  std::vector<uint32_t> ord(nk);
  for (uint32_t k = 0; k < nk; k++) ord[k] = k;
  std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
    return std::memcmp(text.data() + (size_t)K * a, text.data() + (size_t)K * b, (size_t)K) < 0;
  });
  char* ko = (char*)malloc((size_t)K * nk + 1);
  uint8_t* no = (uint8_t*)malloc((size_t)nk + 1);
  uint8_t* ao = (uint8_t*)malloc((size_t)nk + 1);
  if (!ko || !no || !ao) { free(ko); free(no); free(ao); return fail(PL_E_OOM, "host allocation failed"); }
  for (uint32_t i = 0; i < nk; i++) {
    std::memcpy(ko + (size_t)K * i, text.data() + (size_t)K * ord[i], (size_t)K);
    no[i] = nx[ord[i]];
    ao[i] = al[ord[i]];
  }
  *kmers = ko; *next = no; *alt = ao; *n_out = nk;
  if (device_ms) *device_ms = t;
  return PL_OK;
}

}  // namespace

extern "C" {

pl_status pl_kmer_graph(pl_ctx* c, const pl_read_soa* reads, int32_t K, int32_t min_depth, char** kmers,
                        uint8_t** next, uint8_t** alt, int64_t* n, double* device_ms) {
  if (!c || !reads || !kmers || !next || !alt || !n) return fail(PL_E_INVAL, "null argument");
  return kmer_graph(c, reads, K, min_depth, kmers, next, alt, n, device_ms);
}

}  // extern "C"
