// pl_track.hip — the tracks' host side (included at the end of
// pl_engine.hip, after pl_vcf.hip): Tracks.standardTracks values and
// Tracks.makeTrack text for the last batch of a lane, computed by the
// k_gcv_*, k_trk_* and k_wig_lines kernels.

namespace {

const char* const kTrackNames[PL_NTRACKS] = {
    "Changes", "Unconfirmed", "Copy Number", "Coverage", "Bad Coverage", "Pct Bad", "Delta Coverage",
    "Dip Coverage", "Physical Coverage", "Clipped Alignments", "Weighted Qual", "Weighted MQ", "GC"};

// computeGc's buffer values of the listed contigs that have none yet, built
// together in one device block
pl_status build_gcv(pl_ctx* c, const std::vector<int32_t>& ids, hipStream_t st) {
  std::vector<int32_t> miss;
  for (int32_t id : ids)
    if (!c->contigs.at(id).gcv) miss.push_back(id);
  std::sort(miss.begin(), miss.end());
  miss.erase(std::unique(miss.begin(), miss.end()), miss.end());
  if (miss.empty()) return PL_OK;
  std::vector<GcvJob> jobs;
  std::vector<size_t> off;
  size_t bytes = 0;
  int64_t nt = 0;
  for (int32_t id : miss) {
    const Contig& ct = c->contigs.at(id);
    GcvJob g;
    g.ref = ct.dev;
    g.gcv = nullptr;
    g.n = ct.len;
    const int64_t rows = (ct.len + 99) / 100;
    g.rpc = std::max<int64_t>(PL_GCV_MIN_ROWS, (rows + PL_GCV_MAX_CHUNKS - 1) / PL_GCV_MAX_CHUNKS);
    g.chunks = (rows + g.rpc - 1) / g.rpc;
    g.t0 = nt;
    nt += g.chunks * 100;
    off.push_back(bytes);
    bytes += ((size_t)ct.len + 255) & ~(size_t)255;
    jobs.push_back(g);
  }
  uint8_t* block = nullptr;
  if (hipMalloc((void**)&block, bytes) != hipSuccess)
    return fail(PL_E_OOM, "device allocation of %zu bytes failed", bytes);
  std::shared_ptr<uint8_t> mem(block, [](uint8_t* p) { (void)hipFree(p); });
  for (size_t i = 0; i < jobs.size(); i++) jobs[i].gcv = block + off[i];
  VcfBufs B;
  uint8_t* carry;
  GcvJob* dj;
  pl_status s;
  if ((s = B.get(&carry, (size_t)nt)) || (s = B.get(&dj, sizeof(GcvJob) * jobs.size()))) return s;
  HIPCHK(hipMemcpyAsync(dj, jobs.data(), sizeof(GcvJob) * jobs.size(), hipMemcpyHostToDevice, st));
  const int32_t nj = (int32_t)jobs.size();
  const unsigned blocks = (unsigned)((nt + 255) / 256);
  k_gcv_carry<<<dim3(blocks), dim3(256), 0, st>>>(dj, nj, nt, carry);
  k_gcv_top<<<dim3((unsigned)(((int64_t)nj * 100 + 255) / 256)), dim3(256), 0, st>>>(dj, nj, carry);
  k_gcv_fill<<<dim3(blocks), dim3(256), 0, st>>>(dj, nj, nt, carry);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  for (size_t i = 0; i < miss.size(); i++) {
    Contig& ct = c->contigs.at(miss[i]);
    ct.gcv_mem = mem;
    ct.gcv = block + off[i];
  }
  return PL_OK;
}

// The track values of the lane's last batch into T.vals (device), in B's
// buffers; contigs[j] is region j's contig.  ms accumulates the device time.
pl_status tracks_of_lane(pl_ctx* c, Lane* L, const std::vector<int32_t>& contigs, VcfBufs& B, TrackArgs& T,
                         double* ms, bool issues = false) {
  HIPCHK(hipSetDevice(c->device));
  hipStream_t st = L->stream;
  const int64_t G = L->G;
  const int32_t nj = (int32_t)L->hjobs.size();
  std::memset(&T, 0, sizeof T);
  T.A = L->args;
  // the contigs' GC buffers (once per contig) and each region's pointer
  if ((int32_t)contigs.size() != nj) return fail(PL_E_INVAL, "%d contigs for %d regions", (int)contigs.size(), nj);
  if (pl_status s = build_gcv(c, contigs, st)) return s;
  std::vector<const uint8_t*> gp((size_t)nj);
  for (int32_t j = 0; j < nj; j++) gp[(size_t)j] = c->contigs.at(contigs[(size_t)j]).gcv;
  const uint8_t** dgp;
  pl_status s;
  if ((s = B.get(&dgp, sizeof(void*) * (size_t)nj)) || (s = B.get(&T.frag, 4 * (size_t)G)) ||
      (s = B.get(&T.jsum, 8 * (size_t)nj)) || (s = B.get(&T.vals, 4 * (size_t)PL_NTRACKS * (size_t)G)))
    return s;
  if (issues) {
    if ((s = B.get(&T.mask, (size_t)G)) || (s = B.get(&T.jbad, 8 * (size_t)nj)) || (s = B.get(&T.jgood, 8 * (size_t)nj)))
      return s;
    HIPCHK(hipMemsetAsync(T.jbad, 0, 8 * (size_t)nj, st));
    HIPCHK(hipMemsetAsync(T.jgood, 0, 8 * (size_t)nj, st));
  }
  T.gcv = dgp;
  T.jdepth = L->has_jdepth ? L->jdepth.as<int32_t>() : nullptr;
  HIPCHK(hipMemcpyAsync(dgp, gp.data(), sizeof(void*) * (size_t)nj, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(T.jsum, 0, 8 * (size_t)nj, st));
  EventPair ev;
  HIPCHK(ev.create());
  HIPCHK(hipEventRecord(ev.e0, st));
  k_trk_frag<<<dim3((unsigned)L->n_wins), dim3(256), 0, st>>>(T);
  k_trk_values<<<dim3((unsigned)L->n_wins), dim3(256), 0, st>>>(T);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev.e1, st));
  HIPCHK(hipEventSynchronize(ev.e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, ev.e0, ev.e1));
  *ms += t;
  return PL_OK;
}

pl_status tracks_copy(pl_ctx* c, Lane* L, const std::vector<int32_t>& contigs, int32_t* values, int64_t n_values,
                      double* device_ms) {
  const int64_t G = L->G;
  if (values && n_values < (int64_t)PL_NTRACKS * G)
    return fail(PL_E_INVAL, "values holds %lld entries, the batch needs %lld", (long long)n_values,
                (long long)PL_NTRACKS * G);
  VcfBufs B;
  TrackArgs T;
  double ms = 0;
  if (pl_status s = tracks_of_lane(c, L, contigs, B, T, &ms)) return s;
  if (values)
    if (pl_status cs = copy_to_host(L, values, T.vals, 4 * (size_t)PL_NTRACKS * (size_t)G)) return cs;
  if (device_ms) *device_ms = ms;
  return PL_OK;
}

pl_status wig_of_lane(pl_ctx* c, Lane* L, const std::vector<std::string>& rnames, const std::vector<int32_t>& contigs,
                      int32_t track, uint32_t flags, char** text, int64_t* len, double* device_ms) {
  *text = nullptr;
  *len = 0;
  if (track < 0 || track >= PL_NTRACKS) return fail(PL_E_INVAL, "track %d is not in [0, %d)", track, PL_NTRACKS);
  const int64_t G = L->G;
  const int32_t nj = (int32_t)L->hjobs.size();
  hipStream_t st = L->stream;
  VcfBufs B;
  TrackArgs T;
  double ms = 0;
  if (pl_status s = tracks_of_lane(c, L, contigs, B, T, &ms)) return s;
  std::string names;
  std::vector<int64_t> noff(1, 0);
  for (int32_t j = 0; j < nj; j++) {
    names += rnames[(size_t)j];
    noff.push_back((int64_t)names.size());
  }
  char* dnames;
  int64_t* dnoff;
  uint64_t* off;
  pl_status s;
  if ((s = B.get(&dnames, names.size())) || (s = B.get(&dnoff, 8 * noff.size())) || (s = B.get(&T.len, 4 * (size_t)G)) ||
      (s = B.get(&off, 8 * (size_t)(G + 1))))
    return s;
  const int64_t tiles = std::max<int64_t>(1, (G + SCAN_TILE - 1) / SCAN_TILE);
  uint64_t* tsum;
  if ((s = B.get(&tsum, 8 * (size_t)tiles))) return s;
  HIPCHK(hipMemcpyAsync(dnames, names.data(), names.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(dnoff, noff.data(), 8 * noff.size(), hipMemcpyHostToDevice, st));
  T.track = track;
  T.names = dnames;
  T.name_off = dnoff;
  T.off = off;
  EventPair ev;
  HIPCHK(ev.create());
  HIPCHK(hipEventRecord(ev.e0, st));
  k_wig_lines<false><<<dim3((unsigned)L->n_wins), dim3(256), 0, st>>>(T);
  k_scan64_tiles<<<dim3((unsigned)tiles), dim3(256), 0, st>>>(T.len, G, tsum);
  k_scan64_top<<<dim3(1), dim3(1024), 0, st>>>(tsum, tiles);
  k_scan64_apply<<<dim3((unsigned)tiles), dim3(256), 0, st>>>(T.len, G, tsum, off);
  HIPCHK(hipGetLastError());
  uint64_t total = 0;
  HIPCHK(hipMemcpyAsync(&total, off + G, 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  char* dtext;
  if ((s = B.get(&dtext, (size_t)total))) return s;
  T.text = dtext;
  k_wig_lines<true><<<dim3((unsigned)L->n_wins), dim3(256), 0, st>>>(T);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(ev.e1, st));
  std::string head;
  if (flags & PL_WIG_TRACK_LINE) {
    head = std::string("track type=wiggle_0 graphType=line color=0,0,255 altColor=255,0,0 name=\"") +
           kTrackNames[track] + "\"";
    if (track == PL_TRACK_GC) head += " graphType=heatmap midRange=35:65 midColor=0,255,0";
    head += '\n';
  }
  char* out = (char*)malloc(head.size() + (size_t)total + 1);
  if (!out) return fail(PL_E_OOM, "host allocation of %llu bytes failed", (unsigned long long)total);
  std::memcpy(out, head.data(), head.size());
  if (pl_status cs = copy_to_host(L, out + head.size(), dtext, (size_t)total)) {
    free(out);
    return cs;
  }
  out[head.size() + total] = 0;
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, ev.e0, ev.e1));
  if (device_ms) *device_ms = ms + t;
  *text = out;
  *len = (int64_t)(head.size() + total);
  return PL_OK;
}

// GenomeRegion.summaryRegions (GenomeRegion.scala:740-763) over the issue
// bits of one region: runs of loci with the bit, a run ending where more
// than `slop` loci without it follow; runs within 100 loci of either end
// of the region are dropped (nearEdge, :691).  Regions are 1-based loci.
struct Summary {
  uint8_t bit;
  int32_t slop;
  int32_t first = -1, last = -1;
  std::vector<std::pair<int32_t, int32_t>> out;
};

void summary_close(Summary& S, int32_t start, int32_t stop) {
  const int32_t a = start + S.first, b = start + S.last;
  if (!(a - start < 100 || stop - b < 100)) S.out.emplace_back(a, b);
}

// The issue regions of one region from its issue bits: unconfirmed, change,
// low-coverage, collapsed-repeat, duplication, gap and break regions, in
// that order (Tracks.scala:204-210)
enum { ISSUE_GAPS = 5, ISSUE_BREAKS = 6 };
void issue_lists(const uint8_t* m, int32_t start, int32_t stop, std::vector<std::pair<int32_t, int32_t>> lists[7]) {
  const int32_t n = stop + 1 - start;
  Summary S[7] = {{PL_ISSUE_UNCONFIRMED, 100}, {PL_ISSUE_CHANGED, 1}, {PL_ISSUE_LOWCOV, 100}, {PL_ISSUE_HIGHCN, 100},
                  {PL_ISSUE_HIGHCN, 2000}, {PL_ISSUE_N, 100}, {PL_ISSUE_BREAK, 200}};
  for (int32_t i = 0; i < n; i++) {
    while (i + 8 <= n) {                       // skip loci without issues, eight at a time
      uint64_t w;
      std::memcpy(&w, m + i, 8);
      if (w) break;
      i += 8;
    }
    if (i >= n) break;
    const uint8_t b = m[i];
    if (!b) continue;
    for (Summary& q : S) {
      if (!(b & q.bit)) continue;
      if (q.last >= 0 && i > q.last + q.slop + 1) { summary_close(q, start, stop); q.first = -1; }
      if (q.first < 0) q.first = i;
      q.last = i;
    }
  }
  for (Summary& q : S)
    if (q.last >= 0) summary_close(q, start, stop);
  lists[0] = std::move(S[0].out);
  lists[1] = std::move(S[1].out);
  lists[2] = std::move(S[2].out);
  for (auto& r : S[3].out)                     // possibleCollapsedRepeats (:724-729)
    if ((m[r.first - start] & PL_ISSUE_DELTA) && (m[r.second - start] & PL_ISSUE_DELTA)) lists[3].push_back(r);
  for (auto& r : S[4].out)                     // duplicationEvents (:731-737)
    if (r.second + 1 - r.first > 10000) lists[4].push_back(r);
  for (auto& r : S[5].out)                     // gaps (:698)
    if (r.second + 1 - r.first >= 10) lists[5].push_back(r);
  for (auto& r : S[6].out) {                   // possibleBreaks: not near a gap (:720, Region.scala:37-40)
    bool near = false;
    for (auto& g : lists[5])
      if ((r.first <= g.second && r.second >= g.first) || std::abs(r.second - g.first) <= 300 ||
          std::abs(r.first - g.second) <= 300) { near = true; break; }
    if (!near) lists[6].push_back(r);
  }
}

// the region's lines of Tracks.makeBedTrack (Tracks.scala:194-216)
void bed_of_region(const uint8_t* m, int32_t start, int32_t stop, const std::string& name, std::string& out) {
  std::vector<std::pair<int32_t, int32_t>> lists[7];
  issue_lists(m, start, stop, lists);
  static const char* const labels[7] = {"?", "X", "LowCov", "Copy#", "Duplication", "Gap", "Break"};
  static const char* const rgbs[7] = {"0,128,128", "255,0,0", "128,0,128", "128,128,0", "255,128,0", "0,0,0", "255,0,255"};
  for (int k = 0; k < 7; k++)
    for (auto& r : lists[k]) {
      const std::string a = std::to_string(r.first - 1), b = std::to_string(r.second);
      out += name; out += '\t'; out += a; out += '\t'; out += b; out += '\t'; out += labels[k];
      out += "\t0\t+\t"; out += a; out += '\t'; out += b; out += '\t'; out += rgbs[k]; out += '\n';
    }
}

// changed (or null): loci of the batch's only region to count as changed
// besides the device's (those writeVcf marked deleted after a reassembly
// record)
pl_status bed_of_lane(pl_ctx* c, Lane* L, const std::vector<std::string>& rnames, const std::vector<int32_t>& contigs,
                      uint32_t flags, char** text, int64_t* len, double* device_ms,
                      const std::vector<uint8_t>* changed = nullptr) {
  *text = nullptr;
  *len = 0;
  const int64_t G = L->G;
  const int32_t nj = (int32_t)L->hjobs.size();
  VcfBufs B;
  TrackArgs T;
  double ms = 0;
  if (pl_status s = tracks_of_lane(c, L, contigs, B, T, &ms, true)) return s;
  std::vector<uint8_t> mask((size_t)G);
  HIPCHK(hipMemcpyAsync(mask.data(), T.mask, (size_t)G, hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  if (changed && nj == 1 && (int64_t)changed->size() == L->hjobs[0].stop + 1 - L->hjobs[0].start)
    for (size_t i = 0; i < changed->size(); i++)
      if ((*changed)[i]) mask[(size_t)L->hjobs[0].out + i] |= PL_ISSUE_CHANGED;
  // regions in parallel on the host's cores, joined in batch order
  std::vector<std::string> parts((size_t)nj);
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<int32_t> next(0);
  auto work = [&]() {
    for (int32_t j; (j = next.fetch_add(1)) < nj;) {
      const JobDev& J = L->hjobs[(size_t)j];
      bed_of_region(mask.data() + J.out, J.start, J.stop, rnames[(size_t)j], parts[(size_t)j]);
    }
  };
  std::vector<std::thread> th;
  try {
    for (int k = 1; k < std::min(nt, nj); k++) th.emplace_back(work);
  } catch (const std::exception&) {
    // no more threads: the ones started and this one take every region
  }
  work();
  for (auto& t : th) t.join();
  std::string head;
  if (flags & PL_WIG_TRACK_LINE) head = "track description=\"Issues found by Pilon\" name=\"Pilon\"\n";
  size_t total = head.size();
  for (auto& p : parts) total += p.size();
  char* out = (char*)malloc(total + 1);
  if (!out) return fail(PL_E_OOM, "host allocation of %zu bytes failed", total);
  size_t o = 0;
  std::memcpy(out, head.data(), head.size());
  o += head.size();
  for (auto& p : parts) { std::memcpy(out + o, p.data(), p.size()); o += p.size(); }
  out[total] = 0;
  if (device_ms) *device_ms = ms;
  *text = out;
  *len = (int64_t)total;
  return PL_OK;
}

// The pre-Pass-1 depth of region r's "jumps" reads, per position, into the
// main lane (processBam leaves those BAMs out of fragCoverage,
// GenomeRegion.scala:287-300): the jumps alone are piled on the context's
// auxiliary lane and k_trk_frag recovers their depth before Pass 1.
pl_status region_jump_depth(pl_region* r) {
  if (!r->jtemp && r->jtemp_id < 0) return PL_OK;
  pl_ctx* c = r->ctx;
  Lane* L = &c->main;
  Lane* X = &c->aux;
  const int32_t jid = r->jtemp_id;
  c->rsets[jid] = std::move(r->jtemp);
  c->rs_version++;
  const std::vector<JobReq> q{{jid, r->start, r->stop}};
  pl_status s = run_jobs(c, X, q, PL_DEFAULT_WINDOW, 0, PL_DEFAULT_BLOCK, false);
  if (!s) s = fetch_batch_meta(X);
  if (!s) {
    VcfBufs B;
    TrackArgs T;
    std::memset(&T, 0, sizeof T);
    T.A = X->args;
    if (L->G != X->G) s = fail(PL_E_DEVICE, "jumps batch of %lld positions for a region of %lld", (long long)X->G,
                               (long long)L->G);
    else if (L->jdepth.ensure(4 * (size_t)L->G) != hipSuccess) s = fail(PL_E_OOM, "device allocation failed");
    else if (!(s = B.get(&T.jsum, 8 * X->hjobs.size()))) {
      T.frag = L->jdepth.as<int32_t>();
      if (hipMemsetAsync(T.jsum, 0, 8 * X->hjobs.size(), X->stream) != hipSuccess ||
          (k_trk_frag<<<dim3((unsigned)X->n_wins), dim3(256), 0, X->stream>>>(T), hipGetLastError()) != hipSuccess ||
          hipStreamSynchronize(X->stream) != hipSuccess)
        s = fail(PL_E_DEVICE, "jumps depth pass failed");
      else
        L->has_jdepth = true;
    }
  }
  r->jtemp = std::move(c->rsets[jid]);
  c->rs_version++;
  return s;
}

}  // namespace

extern "C" {

pl_status pl_batch_tracks(pl_ctx* c, int32_t* values, int64_t n_values, double* device_ms) {
  if (!c) return fail(PL_E_INVAL, "null argument");
  std::vector<std::string> names;
  std::vector<int32_t> contigs;
  if (pl_status s = batch_names(c, &c->main, &names, &contigs)) return s;
  return tracks_copy(c, &c->main, contigs, values, n_values, device_ms);
}

pl_status pl_region_tracks(pl_region* r, int32_t* values, int64_t n_values, double* device_ms) {
  if (!r) return fail(PL_E_INVAL, "null argument");
  return on_region_lane(r, [&](pl_ctx* c, Lane* L, const std::vector<std::string>&,
                               const std::vector<int32_t>& contigs) {
    return tracks_copy(c, L, contigs, values, n_values, device_ms);
  });
}

pl_status pl_batch_wig(pl_ctx* c, int32_t track, uint32_t flags, char** text, int64_t* len, double* device_ms) {
  if (!c || !text || !len) return fail(PL_E_INVAL, "null argument");
  std::vector<std::string> names;
  std::vector<int32_t> contigs;
  if (pl_status s = batch_names(c, &c->main, &names, &contigs)) return s;
  return wig_of_lane(c, &c->main, names, contigs, track, flags, text, len, device_ms);
}

pl_status pl_region_wig(pl_region* r, int32_t track, uint32_t flags, char** text, int64_t* len, double* device_ms) {
  if (!r || !text || !len) return fail(PL_E_INVAL, "null argument");
  return on_region_lane(r, [&](pl_ctx* c, Lane* L, const std::vector<std::string>& names,
                               const std::vector<int32_t>& contigs) {
    return wig_of_lane(c, L, names, contigs, track, flags, text, len, device_ms);
  });
}

pl_status pl_batch_bed(pl_ctx* c, uint32_t flags, char** text, int64_t* len, double* device_ms) {
  if (!c || !text || !len) return fail(PL_E_INVAL, "null argument");
  std::vector<std::string> names;
  std::vector<int32_t> contigs;
  if (pl_status s = batch_names(c, &c->main, &names, &contigs)) return s;
  return bed_of_lane(c, &c->main, names, contigs, flags, text, len, device_ms);
}

pl_status pl_region_bed(pl_region* r, uint32_t flags, char** text, int64_t* len, double* device_ms) {
  if (!r || !text || !len) return fail(PL_E_INVAL, "null argument");
  pl_status s = on_region_lane(r, [&](pl_ctx* c, Lane* L, const std::vector<std::string>& names,
                                      const std::vector<int32_t>& contigs) {
    return bed_of_lane(c, L, names, contigs, flags, text, len, device_ms, r->vcf_embed.empty() ? nullptr : &r->vcf_embed);
  });
  if (s || r->outcomes.empty()) return s;
  // the reassembly outcomes after the region's issue lines (Tracks.scala:211-213)
  std::string extra;
  const std::string& name = r->ctx->contigs[r->contig].name;
  for (const plh::Outcome& o : r->outcomes) {
    const std::string a = std::to_string(o.start - 1), b = std::to_string(o.stop);
    extra += name + '\t' + a + '\t' + b + '\t' + o.outcome + "\t0\t+\t" + a + '\t' + b + "\t0,128,0\n";
  }
  char* t = (char*)realloc(*text, (size_t)*len + extra.size() + 1);
  if (!t) return fail(PL_E_OOM, "host allocation failed");
  std::memcpy(t + *len, extra.data(), extra.size() + 1);
  *text = t;
  *len += (int64_t)extra.size();
  return PL_OK;
}

}  // extern "C"
