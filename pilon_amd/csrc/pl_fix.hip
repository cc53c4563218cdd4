// pl_fix.hip — the output side of GenomeFile.processRegions (included at the
// end of pl_engine.hip): the .changes lines and the polished FASTA of a whole
// batch, the FASTA element formatter, the region list (contigRegions,
// --targets) and the served requests' fix outputs.  Host code: the change
// lists are sparse (a few per 10 kb), the bases are spliced once per region.
//
// Reference: broadinstitute/pilon v1.24, src/main/scala/org/broadinstitute/pilon/

namespace plh {

// GenomeFile.writeFastaElement (GenomeFile.scala:79-82): ">" header, then the
// sequence in lines of 80 (the last one shorter), each ended by '\n'.  An
// empty sequence gives the header line alone (Scala 2.13's sliding yields no
// group for an empty sequence).
inline void fasta_element(std::string& out, const std::string& header, const char* seq, size_t n) {
  out.reserve(out.size() + header.size() + 2 + n + n / 80 + 1);
  out += '>';
  out += header;
  out += '\n';
  for (size_t i = 0; i < n; i += 80) {
    out.append(seq + i, std::min<size_t>(80, n - i));
    out += '\n';
  }
}

// Java's String.split(",") : trailing empty strings removed, an input with
// no comma (the empty string too) is one piece.
inline std::vector<std::string> java_split_comma(const std::string& s) {
  std::vector<std::string> v;
  size_t a = 0;
  for (;;) {
    const size_t b = s.find(',', a);
    if (b == std::string::npos) { v.push_back(s.substr(a)); break; }
    v.push_back(s.substr(a, b - a));
    a = b + 1;
  }
  if (v.size() > 1)
    while (!v.empty() && v.back().empty()) v.pop_back();
  return v;
}

inline bool java_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == 0x0b || c == '\f' || c == '\r'; }

// The pattern "^\s*(.+):([0-9]+)-([0-9]+)\s*$" of parseTargetString
// (GenomeFile.scala:220-223) with Java's greedy matching: the longest leading
// whitespace first, then the longest name (no line terminator in it) whose
// rest is ":digits-digits" and trailing whitespace.  Returns false for no
// match; *num_ok false when a number leaves Int (Integer.parseInt throws).
inline bool match_target(const std::string& t, std::string* name, int64_t* start, int64_t* stop, bool* num_ok) {
  size_t ws = 0;
  while (ws < t.size() && java_space(t[ws])) ws++;
  for (size_t i = ws + 1; i-- > 0;) {
    // the name s[i, k) may not hold '\n' or '\r' ('.' stops there)
    size_t lim = i;
    while (lim < t.size() && t[lim] != '\n' && t[lim] != '\r') lim++;
    for (size_t k = lim; k > i; k--) {
      if (k >= t.size() || t[k] != ':') continue;
      size_t p = k + 1, d0 = p;
      while (p < t.size() && t[p] >= '0' && t[p] <= '9') p++;
      if (p == d0 || p >= t.size() || t[p] != '-') continue;
      const size_t m = p;
      p++;
      const size_t d1 = p;
      while (p < t.size() && t[p] >= '0' && t[p] <= '9') p++;
      if (p == d1) continue;
      size_t q = p;
      while (q < t.size() && java_space(t[q])) q++;
      if (q != t.size()) continue;
      *name = t.substr(i, k - i);
      auto parse = [num_ok](const std::string& d, int64_t* v) {
        int64_t x = 0;
        for (char ch : d) {
          x = x * 10 + (ch - '0');
          if (x > INT32_MAX) { *num_ok = false; return; }
        }
        *v = x;
      };
      *num_ok = true;
      parse(t.substr(d0, m - d0), start);
      parse(t.substr(d1, p - d1), stop);
      return true;
    }
  }
  return false;
}

}  // namespace plh

namespace {

// .changes lines, fixed bases and counts of job j of the lane's last batch
pl_status job_fix(pl_ctx* c, Lane* L, size_t j, const std::string& new_name, int64_t offset, plh::FixResult* out) {
  std::vector<plh::Change> ch;
  if (pl_status s = job_changes(c, L, j, nullptr, &ch)) return s;
  const JobDev& J = L->hjobs[j];
  const Contig& ct = c->contigs.at(c->rsets[J.rs]->contig);
  const std::string rb((const char*)ct.host.data() + J.start - 1, (size_t)J.size);
  *out = plh::fix_region(c->pc, ct.name, J.start, rb, ch, new_name, offset);
  return PL_OK;
}

char* dup_text(const std::string& s) {
  char* p = (char*)malloc(s.size() + 1);
  if (p) std::memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

}  // namespace

extern "C" {

pl_status pl_batch_fix(pl_ctx* c, const int32_t* element, int32_t n_element, char** changes_text,
                       int64_t* changes_len, char** fasta_text, int64_t* fasta_len) {
  if (!c) return fail(PL_E_INVAL, "null argument");
  Lane* L = &c->main;
  std::vector<std::string> names;
  std::vector<int32_t> contigs;
  if (pl_status s = batch_names(c, L, &names, &contigs)) return s;
  const size_t n = L->hjobs.size();
  if (element && n_element != (int32_t)n)
    return fail(PL_E_INVAL, "element holds %d entries, the last batch has %zu regions", n_element, n);
  for (size_t j = 0; j < n; j++) if (pl_status s = job_status(L, j)) return s;
  // elements: maximal runs of consecutive regions with one element id (the
  // contig when none is given), each one entry of GenomeFile.regions
  std::vector<size_t> el0;                        // first region of each element
  for (size_t j = 0; j < n; j++) {
    const int32_t id = element ? element[j] : contigs[j];
    const int32_t prev = j ? (element ? element[j - 1] : contigs[j - 1]) : 0;
    if (j == 0 || id != prev || contigs[j] != contigs[j - 1]) el0.push_back(j);
  }
  el0.push_back(n);
  const size_t ne = el0.size() - 1;
  // per element: offsets accumulate over its regions (GenomeFile.scala:149-153),
  // so an element's regions are fixed in order by one task; elements in parallel
  std::vector<std::vector<plh::FixResult>> res(ne);
  std::vector<pl_status> err(ne, PL_OK);
  std::vector<std::string> msg(ne);
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t e; (e = next.fetch_add(1)) < ne;) {
      const std::string nn = plh::pilon_name(names[el0[e]]);
      int64_t offset = 0;
      for (size_t j = el0[e]; j < el0[e + 1]; j++) {
        plh::FixResult fr;
        if (pl_status s = job_fix(c, L, j, nn, offset, &fr)) { err[e] = s; msg[e] = g_err; break; }
        offset += (int64_t)fr.bases.size() - L->hjobs[j].size;
        res[e].push_back(std::move(fr));
      }
    }
  };
  const unsigned nth = (unsigned)std::min<size_t>(ne, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nth; t++) {
    try { pool.emplace_back(work); } catch (...) { break; }
  }
  work();
  for (std::thread& t : pool) t.join();
  for (size_t e = 0; e < ne; e++) if (err[e]) { g_err = msg[e]; return err[e]; }
  if (changes_text || changes_len) {
    std::string all;
    for (size_t e = 0; e < ne; e++) for (const plh::FixResult& fr : res[e]) all += fr.changes;
    if (changes_len) *changes_len = (int64_t)all.size();
    if (changes_text && !(*changes_text = dup_text(all))) return fail(PL_E_OOM, "host allocation failed");
  }
  if (fasta_text || fasta_len) {
    std::string fa;
    for (size_t e = 0; e < ne; e++) {
      std::string seq;
      for (const plh::FixResult& fr : res[e]) seq += fr.bases;   // chunk bases concatenated (:156-159)
      plh::fasta_element(fa, plh::pilon_name(names[el0[e]]), seq.data(), seq.size());
    }
    if (fasta_len) *fasta_len = (int64_t)fa.size();
    if (fasta_text && !(*fasta_text = dup_text(fa))) {
      if (changes_text && *changes_text) { free(*changes_text); *changes_text = nullptr; }
      return fail(PL_E_OOM, "host allocation failed");
    }
  }
  return PL_OK;
}

pl_status pl_fasta_element(const char* header, const uint8_t* seq, int64_t len, char** text, int64_t* text_len) {
  if (!header || (!seq && len > 0) || len < 0 || !text) return fail(PL_E_INVAL, "bad argument");
  std::string out;
  plh::fasta_element(out, header, (const char*)seq, (size_t)len);
  if (!(*text = dup_text(out))) return fail(PL_E_OOM, "host allocation failed");
  if (text_len) *text_len = (int64_t)out.size();
  return PL_OK;
}

pl_status pl_pilon_name(const char* name, char** out) {
  if (!name || !out) return fail(PL_E_INVAL, "null argument");
  if (!(*out = dup_text(plh::pilon_name(name)))) return fail(PL_E_OOM, "host allocation failed");
  return PL_OK;
}

pl_status pl_contig_regions(int64_t length, int32_t chunk_size, int32_t* starts, int32_t* stops, int32_t cap,
                            int32_t* n) {
  if (!n || length < 1 || length > INT32_MAX || chunk_size <= 0) return fail(PL_E_INVAL, "bad contig length or chunk size");
  // Int arithmetic as GenomeFile.contigRegions (GenomeFile.scala:67-74)
  const int32_t len = (int32_t)length;
  const int32_t nChunks = j_div(j_add(len, chunk_size - 1), chunk_size);
  if (nChunks <= 0) return fail(PL_E_INVAL, "contig length %d with chunk size %d overflows an Int", len, chunk_size);
  const int32_t chunk = j_div(j_add(len, nChunks - 1), nChunks);
  if (chunk <= 0) return fail(PL_E_INVAL, "chunk size overflows an Int");
  int32_t k = 0;
  for (int64_t base = 1; base < (int64_t)len + 1; base += chunk, k++) {
    if (starts && stops && k < cap) {
      starts[k] = (int32_t)base;
      stops[k] = (int32_t)std::min<int64_t>(len, base + chunk - 1);
    }
  }
  *n = k;
  if ((starts || stops) && k > cap) return fail(PL_E_INVAL, "%d regions, room for %d", k, cap);
  return PL_OK;
}

pl_status pl_parse_targets(const char* targets, const char* const* names, const int64_t* lengths, int32_t n_contigs,
                           pl_target** out, int32_t* n_out) {
  if (!targets || !out || !n_out || (n_contigs > 0 && (!names || !lengths))) return fail(PL_E_INVAL, "null argument");
  std::map<std::string, int32_t> idx;
  for (int32_t i = 0; i < n_contigs; i++) idx.emplace(names[i], i);   // the first contig of a name wins
  auto parse_string = [&](const std::string& s, std::vector<pl_target>& v) -> pl_status {
    for (const std::string& t : plh::java_split_comma(s)) {
      std::string nm;
      int64_t a = 0, b = 0;
      bool num_ok = true;
      const bool m = plh::match_target(t, &nm, &a, &b, &num_ok);
      if (!m) nm = t;                                              // MatchError: the whole contig
      else if (!num_ok) return fail(PL_E_INVAL, "target \"%s\": a coordinate is not an Int", t.c_str());
      auto it = idx.find(nm);
      if (it == idx.end()) return fail(PL_E_INVAL, "target \"%s\": no contig named \"%s\"", t.c_str(), nm.c_str());
      const int64_t len = lengths[it->second];
      if (!m) { a = 1; b = len; }
      // GenomeRegion requires stop <= length; start < 1 or stop < start give
      // regions the rest of Pilon cannot process
      if (b > len) return fail(PL_E_INVAL, "target \"%s\": GenomeRegion stop point must be within contig", t.c_str());
      if (a < 1 || b < a) return fail(PL_E_INVAL, "target \"%s\": start must be in [1, stop]", t.c_str());
      v.push_back(pl_target{it->second, (int32_t)a, (int32_t)b});
    }
    return PL_OK;
  };
  std::vector<pl_target> v;
  // parseTargets (GenomeFile.scala:238-250): the lines of the file named by
  // the string if it reads and parses to something, else the string itself
  bool from_file = false;
  if (FILE* f = fopen(targets, "rb")) {
    std::string text;
    char buf[65536];
    for (size_t k; (k = fread(buf, 1, sizeof buf, f)) > 0;) text.append(buf, k);
    fclose(f);
    std::vector<pl_target> fv;
    bool ok = true;
    size_t a = 0;
    while (a < text.size() && ok) {        // getLines: \n, \r\n or \r ends a line
      size_t b = a;
      while (b < text.size() && text[b] != '\n' && text[b] != '\r') b++;
      const std::string line = text.substr(a, b - a);
      if (b < text.size() && text[b] == '\r' && b + 1 < text.size() && text[b + 1] == '\n') b++;
      a = b + 1;
      if (parse_string(line, fv) != PL_OK) ok = false;
    }
    if (ok && !fv.empty()) { v.swap(fv); from_file = true; }
  }
  if (!from_file) {
    if (pl_status s = parse_string(targets, v)) return s;
  }
  *n_out = (int32_t)v.size();
  *out = (pl_target*)malloc(sizeof(pl_target) * std::max<size_t>(1, v.size()));
  if (!*out) return fail(PL_E_OOM, "host allocation failed");
  if (!v.empty()) std::memcpy(*out, v.data(), sizeof(pl_target) * v.size());
  return PL_OK;
}

pl_status pl_serve_fetch_fix(pl_server* S, int64_t ticket, const char* new_name, int64_t offset, char** changes_text,
                             int64_t* changes_len, uint8_t** bases, int64_t* bases_len) {
  if (!S) return fail(PL_E_INVAL, "null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (!(S->o.copy_out & PL_SERVE_KEEP_FIX)) return fail(PL_E_INVAL, "server was opened without PL_SERVE_KEEP_FIX");
  auto it = S->done.find(ticket);
  if (it == S->done.end() || !it->second.keep_fix)
    return fail(PL_E_INVAL, "no kept fix result for ticket %lld (not done, not polled, or fetched)", (long long)ticket);
  ServeDone d = std::move(it->second);
  it->second.keep_pile = d.keep_pile;
  it->second.pile.swap(d.pile);
  it->second.calls.swap(d.calls);
  it->second.keep_fix = false;
  if (!it->second.keep_pile) S->done.erase(it);
  lk.unlock();
  pl_ctx* c = S->c;
  const Contig& ct = c->contigs.at(c->rsets[d.job.rs]->contig);
  const std::string rb((const char*)ct.host.data() + d.job.start - 1, (size_t)(d.job.stop - d.job.start + 1));
  const std::string nn = new_name ? std::string(new_name) : plh::pilon_name(ct.name);
  const plh::FixResult fr = plh::fix_region(c->pc, ct.name, d.job.start, rb, d.changes, nn, offset);
  if (changes_len) *changes_len = (int64_t)fr.changes.size();
  if (bases_len) *bases_len = (int64_t)fr.bases.size();
  if (changes_text && !(*changes_text = dup_text(fr.changes))) return fail(PL_E_OOM, "host allocation failed");
  if (bases) {
    *bases = (uint8_t*)malloc(fr.bases.size() + 1);
    if (!*bases) {
      if (changes_text) { free(*changes_text); *changes_text = nullptr; }
      return fail(PL_E_OOM, "host allocation failed");
    }
    std::memcpy(*bases, fr.bases.data(), fr.bases.size());
  }
  return PL_OK;
}

}  // extern "C"
