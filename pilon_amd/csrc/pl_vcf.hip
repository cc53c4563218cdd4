// pl_vcf.hip — the VCF writer's host side (included at the end of
// pl_engine.hip): Vcf.writeHeader (Vcf.scala:28-70) and, for the last batch
// of a context, the records of GenomeRegion.writeVcf formatted on the device
// by the k_vcf_* kernels.

namespace {

// device buffers of one VCF pass, freed on every exit
struct VcfBufs {
  std::vector<void*> p;
  ~VcfBufs() { for (void* x : p) (void)hipFree(x); }
  template <class T> pl_status get(T** out, size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, n ? n : 1) != hipSuccess) return fail(PL_E_OOM, "device allocation of %zu bytes failed", n);
    p.push_back(q);
    *out = (T*)q;
    return PL_OK;
  }
};

void add_line(std::string& s, const std::string& l) { s += l; s += '\n'; }

constexpr size_t kStageChunk = (size_t)32 << 20;

// memcpy of large pieces split over host threads that live for one
// hand-off, so the copy out of the staging buffers (and the first-touch page
// faults of a new destination) keep up with the DMA of the next piece
class CopyPool {
 public:
  explicit CopyPool(int n) : nt_(n < 1 ? 1 : n) {
    for (int i = 1; i < nt_; i++) th_.emplace_back([this, i] { run(i); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      quit_ = true;
    }
    cv_.notify_all();
    for (std::thread& t : th_) t.join();
  }
  void copy(char* dst, const char* src, size_t n) {
    if (nt_ == 1 || n < ((size_t)1 << 20)) { std::memcpy(dst, src, n); return; }
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = dst; src_ = src; n_ = n; left_ = nt_ - 1; gen_++;
    }
    cv_.notify_all();
    slice(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return left_ == 0; });
  }

 private:
  void slice(int i) {
    const size_t per = ((n_ + (size_t)nt_ - 1) / (size_t)nt_ + 4095) & ~(size_t)4095;
    const size_t a = std::min(n_, (size_t)i * per), b = std::min(n_, a + per);
    if (b > a) std::memcpy(dst_ + a, src_ + a, b - a);
  }
  void run(int i) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
        if (quit_) return;
        seen = gen_;
      }
      slice(i);
      std::lock_guard<std::mutex> lk(mu_);
      if (--left_ == 0) done_.notify_one();
    }
  }
  int nt_;
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0;
  int left_ = 0;
  uint64_t gen_ = 0;
  bool quit_ = false;
};

// host threads for a copy of n bytes out of the staging buffers
int copy_threads(size_t n) {
  return n < ((size_t)64 << 20) ? 1 : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

// The lane's two pinned staging buffers and their events, created on first use.
pl_status stage_ready(Lane* L) {
  for (int k = 0; k < 2; k++) {
    if (!L->stage[k]) {
      if (hipHostMalloc(&L->stage[k], kStageChunk, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        L->stage[k] = nullptr;
        return fail(PL_E_OOM, "pinned host allocation of %zu bytes failed", kStageChunk);
      }
    }
    if (!L->stage_ev[k]) HIPCHK(hipEventCreate(&L->stage_ev[k]));
  }
  return PL_OK;
}

// Device bytes [src, src + bytes) handed to sink(offset, data, n) in order, a
// chunk at a time, through the lane's two pinned staging buffers: the DMA of
// one chunk overlaps the sink of the one before.  sink returns false to
// stop (PL_E_INVAL).  Allocates nothing after the lane's first hand-off.
template <class Sink>
pl_status drain_to_host(Lane* L, const void* src, size_t bytes, Sink&& sink) {
  if (pl_status s = stage_ready(L)) return s;
  hipStream_t st = L->stream;
  size_t pend_off[2] = {0, 0}, pend_len[2] = {0, 0};
  size_t off = 0;
  bool stop = false;
  for (int k = 0; (off < bytes && !stop) || pend_len[0] || pend_len[1]; k ^= 1) {
    if (pend_len[k]) {                                        // drain this buffer's previous chunk
      HIPCHK(hipEventSynchronize(L->stage_ev[k]));
      if (!stop && !sink(pend_off[k], (const char*)L->stage[k], pend_len[k])) stop = true;
      pend_len[k] = 0;
    }
    if (off < bytes && !stop) {
      const size_t n = std::min(kStageChunk, bytes - off);
      HIPCHK(hipMemcpyAsync(L->stage[k], (const char*)src + off, n, hipMemcpyDeviceToHost, st));
      HIPCHK(hipEventRecord(L->stage_ev[k], st));
      pend_off[k] = off;
      pend_len[k] = n;
      off += n;
    }
  }
  if (stop) return fail(PL_E_INVAL, "the caller's sink stopped the hand-off");
  return PL_OK;
}

// A device-to-host copy into pageable memory: small copies directly, large
// ones through the lane's staging buffers.
pl_status copy_to_host(Lane* L, void* dst, const void* src, size_t bytes) {
  if (bytes <= 2 * kStageChunk) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    return PL_OK;
  }
  CopyPool pool(copy_threads(bytes));
  return drain_to_host(L, src, bytes, [dst, &pool](size_t off, const char* p, size_t n) {
    pool.copy((char*)dst + off, p, n);
    return true;
  });
}

// two HIP events bracketing a pass, destroyed on every exit
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  hipError_t create() {
    hipError_t e = hipEventCreate(&e0);
    return e != hipSuccess ? e : hipEventCreate(&e1);
  }
};

}  // namespace

extern "C" {

pl_status pl_vcf_header(const char* const* names, const int64_t* lengths, int32_t n, uint32_t flags,
                        const char* date, const char* source, const char* command, const char* reference,
                        char** text, int64_t* len) {
  if (!text || !len || (n > 0 && (!names || !lengths))) return fail(PL_E_INVAL, "null argument");
  std::string s;
  add_line(s, "##fileformat=VCFv4.1");
  add_line(s, std::string("##fileDate=") + (date ? date : ""));
  add_line(s, std::string("##source=\"") + (source ? source : "") + "\"");
  add_line(s, std::string("##PILON=\"") + (command ? command : "") + "\"");
  add_line(s, std::string("##reference=") + (reference ? reference : ""));
  for (int32_t i = 0; i < n; i++)
    add_line(s, std::string("##contig=<ID=") + names[i] + ",length=" + std::to_string(lengths[i]) + ">");
  add_line(s, "##FILTER=<ID=LowCov,Description=\"Low Coverage of good reads at location\">");
  add_line(s, "##FILTER=<ID=Amb,Description=\"Ambiguous evidence in haploid genome\">");
  add_line(s, "##FILTER=<ID=Del,Description=\"This base is in a deletion or change event from another record\">");
  add_line(s, "##INFO=<ID=DP,Number=1,Type=Integer,Description=\"Valid read depth; some reads may have been filtered\">");
  add_line(s, "##INFO=<ID=TD,Number=1,Type=Integer,Description=\"Total read depth including bad pairs\">");
  add_line(s, "##INFO=<ID=PC,Number=1,Type=Integer,Description=\"Physical coverage of valid inserts across locus\">");
  add_line(s, "##INFO=<ID=BQ,Number=1,Type=Integer,Description=\"Mean base quality at locus\">");
  add_line(s, "##INFO=<ID=MQ,Number=1,Type=Integer,Description=\"Mean read mapping quality at locus\">");
  add_line(s, "##INFO=<ID=QD,Number=1,Type=Integer,Description=\"Variant confidence/quality by depth\">");
  add_line(s, "##INFO=<ID=BC,Number=4,Type=Integer,Description=\"Count of As, Cs, Gs, Ts at locus\">");
  if (flags & PL_VCF_QE)
    add_line(s, "##INFO=<ID=QE,Number=4,Type=Integer,Description=\"Evidence for As, Cs, Gs, Ts weighted by Q & MQ at locus\">");
  else
    add_line(s, "##INFO=<ID=QP,Number=4,Type=Integer,Description=\"Percentage of As, Cs, Gs, Ts weighted by Q & MQ at locus\">");
  add_line(s, "##INFO=<ID=IC,Number=1,Type=Integer,Description=\"Number of reads with insertion here\">");
  add_line(s, "##INFO=<ID=DC,Number=1,Type=Integer,Description=\"Number of reads with deletion here\">");
  add_line(s, "##INFO=<ID=XC,Number=1,Type=Integer,Description=\"Number of reads clipped here\">");
  add_line(s, "##INFO=<ID=AC,Number=A,Type=Integer,Description=\"Allele count in genotypes, for each ALT allele, in the same order as listed\">");
  add_line(s, "##INFO=<ID=AF,Number=A,Type=Float,Description=\"Fraction of evidence in support of alternate allele(s)\">");
  add_line(s, "##INFO=<ID=SVTYPE,Number=1,Type=String,Description=\"Type of structural variant\">");
  add_line(s, "##INFO=<ID=SVLEN,Number=.,Type=String,Description=\"Difference in length between REF and ALT alleles\">");
  add_line(s, "##INFO=<ID=END,Number=1,Type=Integer,Description=\"End position of the variant described in this record\">");
  add_line(s, "##INFO=<ID=IMPRECISE,Number=0,Type=Flag,Description=\"Imprecise change from local reassembly (ALT contains Ns)\">");
  add_line(s, "##FORMAT=<ID=GT,Number=1,Type=String,Description=\"Genotype\">");
  add_line(s, "##FORMAT=<ID=AD,Number=.,Type=String,Description=\"Allelic depths for the ref and alt alleles in the order listed\">");
  add_line(s, "##FORMAT=<ID=DP,Number=1,Type=String,Description=\"Approximate read depth; some reads may have been filtered\">");
  add_line(s, "##ALT=<ID=DUP,Description=\"Possible segmental duplication\">");
  add_line(s, "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\tSAMPLE");
  char* out = (char*)malloc(s.size() + 1);
  if (!out) return fail(PL_E_OOM, "host allocation failed");
  std::memcpy(out, s.data(), s.size() + 1);
  *text = out;
  *len = (int64_t)s.size();
  return PL_OK;
}

}  // extern "C"

namespace {

pl_status tracks_of_lane(pl_ctx* c, Lane* L, const std::vector<int32_t>& contigs, VcfBufs& B, TrackArgs& T,
                         double* ms, bool issues);

// GenomeRegion.duplicationEvents (GenomeRegion.scala:731-737) of one region
// from its issue bits: summaryRegions of copyNumber > 1 with slop 2000
// (:740-763, runs within 100 loci of an end dropped by nearEdge), longer
// than 10,000 loci.  1-based (start, stop) pairs in order.
std::vector<std::pair<int32_t, int32_t>> dup_runs(const uint8_t* m, int32_t start, int32_t stop) {
  std::vector<std::pair<int32_t, int32_t>> out;
  const int32_t n = stop + 1 - start;
  int32_t first = -1, last = -1;
  auto close = [&]() {
    const int32_t a = start + first, b = start + last;
    if (!(a - start < 100 || stop - b < 100) && b + 1 - a > 10000) out.emplace_back(a, b);
  };
  for (int32_t i = 0; i < n; i++) {
    if (!(m[i] & PL_ISSUE_HIGHCN)) continue;
    if (last >= 0 && i > last + 2000 + 1) { close(); first = -1; }
    if (first < 0) first = i;
    last = i;
  }
  if (last >= 0) close();
  return out;
}

// A batch's VCF text: the records formatted on the device, and the lines
// the host puts between them (Vcf.writeDup before a duplication event's
// first locus) at their byte offsets in the device text.
struct VcfText {
  VcfBufs B;
  const char* dtext = nullptr;
  uint64_t total = 0;
  std::vector<std::pair<uint64_t, std::string>> inserts;   // ascending offsets
  double device_ms = 0;
  uint64_t bytes() const {
    uint64_t n = total;
    for (const auto& x : inserts) n += x.second.size();
    return n;
  }
};

// The text in order to sink(data, n): device pieces through the lane's
// staging buffers, the inserted lines between them.
template <class Sink>
pl_status vcf_emit(Lane* L, const VcfText& V, Sink&& sink) {
  uint64_t at = 0;
  auto piece = [&](uint64_t end) -> pl_status {
    if (end <= at) return PL_OK;
    pl_status s = drain_to_host(L, V.dtext + at, (size_t)(end - at),
                                [&](size_t, const char* p, size_t n) { return sink(p, n); });
    at = end;
    return s;
  };
  for (const auto& x : V.inserts) {
    if (pl_status s = piece(x.first)) return s;
    if (!sink(x.second.data(), x.second.size())) return fail(PL_E_INVAL, "the caller's sink stopped the hand-off");
  }
  return piece(V.total);
}

// A region's local reassembly in its VCF (GenomeRegion.writeVcf,
// GenomeRegion.scala:623-643): the Vcf.writeFixRecord line of each record
// at its locus (after a writeDup line there), and the loci the fixes
// replace, embedded from then on.
struct FixRecords {
  std::vector<uint8_t> embed;                              // [region size]
  std::vector<std::pair<int32_t, std::string>> lines;      // (locus, line), ascending
};

// The VCF records of the lane's last batch; names[j] is region j's name;
// fr (or null) the reassembly records of the batch's only region.
pl_status vcf_format(pl_ctx* c, Lane* L, const std::vector<std::string>& rnames, const std::vector<int32_t>& contigs,
                     uint32_t flags, VcfText& out, const FixRecords* fr = nullptr) {
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = L->stream;
  const int64_t G = L->G;
  const int32_t nj = (int32_t)L->hjobs.size();
  std::string names;
  std::vector<int64_t> noff(1, 0);
  for (int32_t j = 0; j < nj; j++) {
    names += rnames[(size_t)j];
    noff.push_back((int64_t)names.size());
  }
  VcfBufs& B = out.B;
  VcfArgs V;
  std::memset(&V, 0, sizeof V);
  V.A = L->args;
  V.flags = flags;
  char* dnames;
  int64_t* dnoff;
  pl_status s;
  if ((s = B.get(&dnames, names.size())) || (s = B.get(&dnoff, 8 * noff.size()))) return s;
  HIPCHK(hipMemcpyAsync(dnames, names.data(), names.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dnoff, noff.data(), 8 * noff.size(), hipMemcpyHostToDevice, st));
  V.names = dnames;
  V.name_off = dnoff;
  if (fr) {
    if (nj != 1 || (int64_t)fr->embed.size() != L->hjobs[0].stop + 1 - L->hjobs[0].start)
      return fail(PL_E_INVAL, "reassembly records for a batch that is not their region");
    uint8_t* emb;
    if ((s = B.get(&emb, (size_t)G))) return s;
    HIPCHK(hipMemsetAsync(emb, 0, (size_t)G, st));
    HIPCHK(hipMemcpyAsync(emb + L->hjobs[0].out, fr->embed.data(), fr->embed.size(), hipMemcpyHostToDevice, st));
    V.embed = emb;
  }
  if ((s = B.get(&V.ilist_n, 4)) || (s = B.get(&V.imap, 4 * (size_t)G)) || (s = B.get(&V.len, 4 * (size_t)G)))
    return s;
  uint64_t* off;
  if ((s = B.get(&off, 8 * (size_t)(G + 1)))) return s;
  V.off = off;
  EventPair ev;
  HIPCHK(ev.create());
  HIPCHK(hipEventRecord(ev.e0, st));
  // indel-call columns, grown once if the first capacity was short
  uint32_t icap = (uint32_t)std::min<int64_t>(G, 1 << 20) + 64, ni = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    if ((s = B.get(&V.ilist, 4 * (size_t)icap)) || (s = B.get(&V.irec, 8 * (size_t)icap))) return s;
    V.icap = icap;
    HIPCHK(hipMemsetAsync(V.ilist_n, 0, 4, st));
    k_vcf_mark<<<dim3((unsigned)L->n_wins), dim3(256), 0, st>>>(V);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&ni, V.ilist_n, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (ni <= icap) break;
    icap = ni + 64;
  }
  if (ni) {
    k_vcf_indel<<<dim3((unsigned)std::min<uint32_t>(4096, (ni + 3) / 4)), dim3(256), 0, st>>>(V);
    HIPCHK(hipGetLastError());
  }
  k_vcf_lines<false><<<dim3((unsigned)L->n_wins), dim3(256), 0, st>>>(V);
  HIPCHK(hipGetLastError());
  const int64_t tiles = std::max<int64_t>(1, (G + SCAN_TILE - 1) / SCAN_TILE);
  uint64_t* tsum;
  if ((s = B.get(&tsum, 8 * (size_t)tiles))) return s;
  k_scan64_tiles<<<dim3((unsigned)tiles), dim3(256), 0, st>>>(V.len, G, tsum);
  k_scan64_top<<<dim3(1), dim3(1024), 0, st>>>(tsum, tiles);
  k_scan64_apply<<<dim3((unsigned)tiles), dim3(256), 0, st>>>(V.len, G, tsum, off);
  HIPCHK(hipGetLastError());
  uint64_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, off + G, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  char* dtext;
  if ((s = B.get(&dtext, (size_t)total))) return s;
  V.text = dtext;
  k_vcf_write<<<dim3((unsigned)L->n_wins), dim3(256), 0, st>>>(V);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev.e1, st));
  HIPCHK(hipEventSynchronize(ev.e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, ev.e0, ev.e1));
  out.dtext = dtext;
  out.total = total;
  out.device_ms = ms;
  // duplication events from Pass 2's copy number (the tracks' issue bits):
  // a writeDup line (Vcf.scala:194-202) before the records of the first
  // locus of each, at that locus's offset in the device text
  {
    VcfBufs B2;
    TrackArgs T;
    double tms = 0;
    if ((s = tracks_of_lane(c, L, contigs, B2, T, &tms, true))) return s;
    std::vector<uint8_t> mask((size_t)G);
    HIPCHK(hipMemcpyAsync(mask.data(), T.mask, (size_t)G, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    out.device_ms += tms;
    for (int32_t j = 0; j < nj; j++) {
      const JobDev& J = L->hjobs[(size_t)j];
      for (const auto& d : dup_runs(mask.data() + J.out, J.start, J.stop)) {
        uint64_t at = 0;
        HIPCHK(hipMemcpyAsync(&at, off + J.out + (d.first - J.start), 8, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        const Contig& ct = c->contigs.at(contigs[(size_t)j]);
        const int32_t loc = d.first - 1;
        std::string line = rnames[(size_t)j] + "\t" + std::to_string(loc) + "\t.\t";
        line += (char)to_upper(ct.host[(size_t)loc - 1]);
        line += "\t<DUP>\t.\tPASS\tSVTYPE=DUP;SVLEN=" + std::to_string(d.second + 1 - d.first) + ";END=" +
                std::to_string(d.second) + ";IMPRECISE\tGT\t./.\n";
        out.inserts.emplace_back(at, std::move(line));
      }
    }
  }
  if (fr && !fr->lines.empty()) {
    const JobDev& J = L->hjobs[0];
    for (const auto& x : fr->lines) {
      uint64_t at = 0;
      HIPCHK(hipMemcpyAsync(&at, off + J.out + (x.first - J.start), 8, hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      out.inserts.emplace_back(at, x.second);
    }
    // a locus's writeDup line comes before its reassembly record
    std::stable_sort(out.inserts.begin(), out.inserts.end(),
                     [](const std::pair<uint64_t, std::string>& a, const std::pair<uint64_t, std::string>& b) {
                       return a.first < b.first;
                     });
  }
  return PL_OK;
}

pl_status vcf_of_lane(pl_ctx* c, Lane* L, const std::vector<std::string>& rnames, const std::vector<int32_t>& contigs,
                      uint32_t flags, char** text, int64_t* len, double* device_ms, const FixRecords* fr = nullptr) {
  *text = nullptr;
  *len = 0;
  VcfText V;
  if (pl_status s = vcf_format(c, L, rnames, contigs, flags, V, fr)) return s;
  const uint64_t n = V.bytes();
  char* out = (char*)malloc((size_t)n + 1);
  if (!out) return fail(PL_E_OOM, "host allocation of %llu bytes failed", (unsigned long long)n);
  uint64_t w = 0;
  CopyPool pool(copy_threads((size_t)n));
  pl_status s = vcf_emit(L, V, [&](const char* p, size_t k) {
    pool.copy(out + w, p, k);
    w += k;
    return true;
  });
  if (s) { free(out); return s; }
  out[n] = 0;
  if (device_ms) *device_ms = V.device_ms;
  *text = out;
  *len = (int64_t)n;
  return PL_OK;
}

// f(ctx, lane, region names, region contigs) on a lane whose last batch is the finished
// region r: the context's own lane if it still holds r's batch, else r's
// batch runs there again (its read set goes back into the context's table
// for the run).
template <class F>
pl_status on_region_lane(pl_region* r, F&& f) {
  if (!r->finished) return fail(PL_E_INVAL, "region not finished (pl_region_finish first)");
  pl_ctx* c = r->ctx;
  Lane* L = &c->main;
  const int32_t id = r->readset >= 0 ? r->readset : r->temp_id;
  const std::vector<JobReq> q{{id, r->start, r->stop}};
  const bool current = L->tables_valid && L->last_reqs == q && L->hjobs.size() == 1;
  std::vector<std::string> names{c->contigs.at(r->contig).name};
  const std::vector<int32_t> contigs{r->contig};
  if (current) return f(c, L, names, contigs);
  HIPCHK(hipSetDevice(c->device));
  if (r->temp) { c->rsets[id] = std::move(r->temp); c->rs_version++; }
  pl_status s = run_jobs(c, L, q, PL_DEFAULT_WINDOW, 0, PL_DEFAULT_BLOCK, false);
  if (!s) s = fetch_batch_meta(L);
  if (!s) s = region_jump_depth(r);
  if (!s) s = f(c, L, names, contigs);
  if (r->temp_id >= 0 && !r->temp) { r->temp = std::move(c->rsets[id]); c->rs_version++; }
  return s;
}

// The reassembly records of a region's VCF: at each kept fix (fixFixList
// over snpFixList ++ smallFixList ++ bigFixList) that is one of the big
// fixes and not a small one, Vcf.writeFixRecord (Vcf.scala:178-192) and the
// loci it replaces marked (GenomeRegion.scala:632-636).  PL_E_EDGE where
// refBase(locus - 1) leaves the region, as Pilon throws there.
pl_status region_fix_records(pl_region* r, FixRecords* out) {
  out->embed.assign((size_t)(r->stop + 1 - r->start), 0);
  out->lines.clear();
  if (r->big.empty()) return PL_OK;
  const Contig& ct = r->ctx->contigs[r->contig];
  std::vector<plh::Fix> snp, small;
  plh::FixResult counts;
  plh::fix_lists(r->ctx->pc, r->changes, snp, small, counts);
  const std::vector<plh::Fix> big(r->big.rbegin(), r->big.rend());
  std::vector<plh::Fix> all = snp;
  all.insert(all.end(), small.begin(), small.end());
  all.insert(all.end(), big.begin(), big.end());
  auto same = [](const plh::Fix& a, const plh::Fix& b) { return a.loc == b.loc && a.was == b.was && a.patch == b.patch; };
  auto in = [&](const std::vector<plh::Fix>& l, const plh::Fix& f) {
    for (const plh::Fix& x : l) if (same(x, f)) return true;
    return false;
  };
  for (const plh::Fix& f : plh::fix_fix_list(all)) {
    if (!in(big, f) || in(small, f)) continue;
    const int32_t loc = f.loc - 1;
    if (loc < r->start) return fail(PL_E_EDGE, "Pilon 1.24 throws here: a reassembly record at the region's first locus");
    const char rb = (char)to_upper(ct.host[(size_t)loc - 1]);
    const std::string ref = rb + f.was, alt = rb + f.patch;
    const int64_t svlen = (int64_t)alt.size() - (int64_t)ref.size();
    std::string line = ct.name + "\t" + std::to_string(loc) + "\t.\t" + ref + "\t" + alt + "\t.\tPASS\tSVTYPE=" +
                       (svlen < 0 ? "DEL" : "INS") + ";SVLEN=" + std::to_string(svlen) + ";END=" +
                       std::to_string((int64_t)loc + (int64_t)ref.size() - 1);
    if (f.patch.find('N') != std::string::npos) line += ";IMPRECISE";
    line += "\tGT\t1/1\n";
    out->lines.emplace_back(f.loc, std::move(line));
    for (size_t j = 0; j < f.was.size(); j++) out->embed[(size_t)(f.loc - r->start) + j] = 1;
  }
  return PL_OK;
}

// the region names and contigs of the lane's last batch (PL_E_INVAL when a
// finished region's reads were released)
pl_status batch_names(pl_ctx* c, Lane* L, std::vector<std::string>* names, std::vector<int32_t>* contigs = nullptr) {
  if (!L->tables_valid || L->hjobs.empty() || L->n_wins <= 0) return fail(PL_E_INVAL, "no batch on the device");
  for (const JobDev& J : L->hjobs) {
    if (J.rs < 0 || J.rs >= (int32_t)c->rsets.size() || !c->rsets[J.rs])
      return fail(PL_E_INVAL, "the last batch's reads were released (a finished region: use the pl_region_* form)");
    names->push_back(c->contigs.at(c->rsets[J.rs]->contig).name);
    if (contigs) contigs->push_back(c->rsets[J.rs]->contig);
  }
  return PL_OK;
}

}  // namespace

extern "C" {

pl_status pl_batch_vcf(pl_ctx* c, uint32_t flags, char** text, int64_t* len, double* device_ms) {
  if (!c || !text || !len) return fail(PL_E_INVAL, "null argument");
  std::vector<std::string> names;
  std::vector<int32_t> contigs;
  if (pl_status s = batch_names(c, &c->main, &names, &contigs)) return s;
  return vcf_of_lane(c, &c->main, names, contigs, flags, text, len, device_ms);
}

pl_status pl_batch_vcf_stream(pl_ctx* c, uint32_t flags, pl_text_sink sink, void* user, int64_t* len,
                              double* device_ms) {
  if (!c || !sink) return fail(PL_E_INVAL, "null argument");
  std::vector<std::string> names;
  std::vector<int32_t> contigs;
  if (pl_status s = batch_names(c, &c->main, &names, &contigs)) return s;
  VcfText V;
  if (pl_status s = vcf_format(c, &c->main, names, contigs, flags, V)) return s;
  if (pl_status s = vcf_emit(&c->main, V, [&](const char* p, size_t n) { return sink(user, p, (int64_t)n) == 0; }))
    return s;
  if (len) *len = (int64_t)V.bytes();
  if (device_ms) *device_ms = V.device_ms;
  return PL_OK;
}

pl_status pl_region_vcf(pl_region* r, uint32_t flags, char** text, int64_t* len, double* device_ms) {
  if (!r || !text || !len) return fail(PL_E_INVAL, "null argument");
  FixRecords fr;
  if (pl_status s = region_fix_records(r, &fr)) return s;
  pl_status s = on_region_lane(r, [&](pl_ctx* c, Lane* L, const std::vector<std::string>& names,
                                      const std::vector<int32_t>& contigs) {
    return vcf_of_lane(c, L, names, contigs, flags, text, len, device_ms, r->big.empty() ? nullptr : &fr);
  });
  if (!s && !r->big.empty()) r->vcf_embed = std::move(fr.embed);
  return s;
}

}  // extern "C"
