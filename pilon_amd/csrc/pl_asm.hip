// pl_asm.hip — local reassembly across a gap or a break (included at the end
// of pl_engine.hip).  The K-mer graph of the recruited reads is built on the
// device (kmer_graph: Assembler.addReads + buildGraph); on it the host runs
// Assembler's path search (Assembler.scala:163-280) and GapFiller's
// assembly across the break (GapFiller.scala:39-262).  Both are sequential
// walks over a few thousand K-mers per break.
//
// A path is kept oldest K-mer first; Pilon keeps it as a List with the newest
// K-mer at the head, so "index from the head" is the distance from the end.

namespace {

struct AsmPath {
  std::vector<const std::string*> kmers;          // oldest first, pointing into the graph's keys
  std::unordered_map<const std::string*, int32_t> count, last;   // occurrences, newest position
  void push(const std::string* k) {
    last[k] = (int32_t)kmers.size();
    count[k]++;
    kmers.push_back(k);
  }
  // List.indexOf from the head: -1 if absent
  int32_t index_of(const std::string* k) const {
    auto it = last.find(k);
    return it == last.end() ? -1 : (int32_t)kmers.size() - 1 - it->second;
  }
  // kmers.tail contains k: an occurrence other than the newest K-mer itself
  bool tail_contains(const std::string* k) const {
    auto it = count.find(k);
    if (it == count.end()) return false;
    return it->second - (kmers.back() == k ? 1 : 0) > 0;
  }
  int32_t occurrences(const std::string* k) const {
    auto it = count.find(k);
    return it == count.end() ? 0 : it->second;
  }
};

struct Assembler {
  int32_t K;
  bool tr_safe;
  // kGraph / altGraph (Assembler.scala:37-39): a K-mer to its successor.
  // Successors are interned so a path holds pointers.
  std::unordered_map<std::string, const std::string*> kg, alt;
  std::unordered_set<std::string> pool;
  int32_t loop_length = 0;
  std::string loop_sequence;

  const std::string* intern(const std::string& s) { return &*pool.insert(s).first; }

  // kmerPathString of the path's newest n K-mers (Assembler.scala:163-170)
  static std::string path_string(const AsmPath& p, size_t n) {
    const size_t b = p.kmers.size() - n;
    std::string s = *p.kmers[b];
    for (size_t i = b + 1; i < p.kmers.size(); i++) s.push_back(p.kmers[i]->back());
    return s;
  }
  void note_loop(int32_t loop_index, const AsmPath& p) {       // Assembler.scala:172-180
    const int32_t length = loop_index + 1;
    if (loop_length == 0 || length < loop_length) {
      loop_length = length;
      loop_sequence = path_string(p, (size_t)length).substr(0, (size_t)length);
    }
  }
  // Assembler.scala:182-233
  void paths_forward_from(AsmPath p, int32_t branches, std::vector<AsmPath>* out) {
    for (;;) {
      const std::string* kmer = p.kmers.back();
      auto it = kg.find(*kmer);
      if (it == kg.end()) { out->push_back(std::move(p)); return; }
      auto ia = alt.find(*kmer);
      if (ia != alt.end()) {
        const std::string* n1 = it->second;
        const std::string* n2 = ia->second;
        const bool seen1 = p.tail_contains(n1), seen2 = p.tail_contains(n2);
        if (seen1 || seen2) {
          note_loop(std::max(p.index_of(n1), p.index_of(n2)), p);
          if (tr_safe) { out->push_back(std::move(p)); return; }
        }
        if (seen1 && seen2) { out->push_back(std::move(p)); return; }
        if (seen1 && !seen2) p.push(n2);
        else if (seen2 && !seen1) p.push(n1);
        else {
          if (branches < 5) {                                    // Assembler.maxBranches
            AsmPath q = p;
            p.push(n1);
            paths_forward_from(std::move(p), branches + 1, out);
            q.push(n2);
            paths_forward_from(std::move(q), branches + 1, out);
          } else {
            out->push_back(std::move(p));
          }
          return;
        }
      } else {
        const std::string* n = it->second;
        const int32_t loop = p.index_of(n);
        if (loop >= 0) {
          note_loop(loop, p);
          if (tr_safe) { out->push_back(std::move(p)); return; }
          if (p.occurrences(n) > 1) { out->push_back(std::move(p)); return; }
        }
        p.push(n);
      }
    }
  }
  // pathsForward (Assembler.scala:235-244): longest first, stable
  std::vector<std::string> paths_forward(const std::string& start) {
    AsmPath p;
    p.push(intern(start));
    std::vector<AsmPath> ps;
    paths_forward_from(std::move(p), 0, &ps);
    std::vector<std::string> out;
    for (const AsmPath& q : ps) out.push_back(path_string(q, q.kmers.size()));
    std::stable_sort(out.begin(), out.end(), [](const std::string& a, const std::string& b) { return a.size() > b.size(); });
    return out;
  }
  // tryForward (Assembler.scala:251-264)
  std::vector<std::string> try_forward(const std::string& anchor) {
    if ((int32_t)anchor.size() < K) return {anchor};
    const std::string starting = anchor.substr(0, (size_t)K);
    std::vector<std::string> paths = paths_forward(starting);
    if (!paths.empty() && paths[0].size() > anchor.size() + 20) return paths;   // minExtend
    std::vector<std::string> rest = try_forward(anchor.substr((size_t)K));
    for (std::string& s : rest) s = starting + s;
    return rest;
  }
};

char asm_complement(char c) {                                   // Bases.complement (Bases.scala:47-53)
  return c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : c == 'T' ? 'A' : c;
}
std::string asm_rc(const std::string& s) {                      // Bases.reverseComplement
  std::string r(s.rbegin(), s.rend());
  for (char& c : r) c = asm_complement(c);
  return r;
}
char asm_upper(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 32) : c; }

struct AsmFix { int32_t start = 0; std::string ref, patch; };
bool asm_none(const AsmFix& f) { return f.start == 0 && f.ref.empty() && f.patch.empty(); }
bool operator==(const AsmFix& a, const AsmFix& b) { return a.start == b.start && a.ref == b.ref && a.patch == b.patch; }

// what GapFiller reads of its GenomeRegion (GenomeRegion.scala:766-781); a
// require that fails in Pilon throws there: here it sets edge
struct AsmRegion {
  int32_t start, stop;
  const char* bases;          // the region's current bases (SNP fixes applied)
  const char* contig;         // the contig as read
  bool edge = false;
  bool in(int64_t l) const { return l >= start && l <= stop; }
  char base_at(int32_t l) { if (!in(l)) { edge = true; return 0; } return asm_upper(bases[l - start]); }
  char original_base_at(int32_t l) { if (!in(l)) { edge = true; return 0; } return asm_upper(contig[l - 1]); }
  std::string sub_string(int32_t l, int32_t n) {
    if (!in(l) || !in((int64_t)l + n - 1)) { edge = true; return std::string(); }
    std::string s(bases + (l - start), (size_t)n);
    for (char& c : s) c = asm_upper(c);
    return s;
  }
  std::string ref_sub_string(int32_t l, int32_t n) {
    if (!in(l) || !in((int64_t)l + n - 1)) { edge = true; return std::string(); }
    std::string s(contig + (l - 1), (size_t)n);
    for (char& c : s) c = asm_upper(c);
    return s;
  }
};

struct GapFill {
  AsmRegion* R;
  Assembler* A;
  pl_asm_opts o;

  AsmFix trim_patch(int32_t start, std::string patch, int32_t stop) {     // GapFiller.scala:239-253
    size_t a = 0, b = patch.size();
    while (start < stop && a < b && (R->base_at(start) == patch[a] || R->original_base_at(start) == patch[a])) {
      start++;
      a++;
    }
    while (start < stop && a < b &&
           (R->base_at(stop - 1) == patch[b - 1] || R->original_base_at(stop - 1) == patch[b - 1])) {
      stop--;
      b--;
    }
    AsmFix f;
    f.start = start;
    f.ref = R->ref_sub_string(start, stop - start);
    f.patch = patch.substr(a, b - a);
    return f;
  }
  static std::string proper_overlap(const std::string& l, const std::string& r, int32_t min_overlap) {  // :214-237
    const int64_t ll = (int64_t)l.size(), rl = (int64_t)r.size();
    for (int64_t ov = min_overlap; ov <= ll + rl - 2 * (int64_t)min_overlap; ov++) {
      const int64_t lo = std::max<int64_t>(ll - ov, 0), ro = std::max<int64_t>(ov - ll, 0);
      const int64_t n = std::min(ll - lo, rl - ro);
      // an empty comparison matches, as substrMatch's empty loop does
      if (n <= 0 || l.compare((size_t)lo, (size_t)n, r, (size_t)ro, (size_t)n) == 0)
        return l.substr(0, (size_t)lo) + r.substr((size_t)ro);
    }
    return std::string();
  }
  // breakJoins (:160-181) with joinBreak (:183-212): the distinct joins in the
  // order found (Pilon's Set lists up to four in insertion order)
  std::vector<AsmFix> break_joins(int32_t start, const std::vector<std::string>& fw, const std::vector<std::string>& rv,
                                  int32_t stop) {
    std::vector<AsmFix> found;
    for (const std::string& f : fw)
      for (const std::string& r : rv) {
        const std::string patch = proper_overlap(f, r, 2 * o.k + 1);      // GapFiller.k
        if (patch.empty()) continue;
        AsmFix s = trim_patch(start, patch, stop);
        if (R->edge) return {};
        if (!asm_none(s) && std::find(found.begin(), found.end(), s) == found.end()) found.push_back(s);
      }
    std::vector<AsmFix> sols = found;
    std::stable_sort(sols.begin(), sols.end(), [](const AsmFix& a, const AsmFix& b) {
      return a.patch.size() + a.ref.size() < b.patch.size() + b.ref.size();
    });
    std::set<int64_t> lengths;
    for (const AsmFix& s : found) lengths.insert((int64_t)s.patch.size() - (int64_t)s.ref.size());
    if (lengths.size() == 1) sols.resize(1);
    return sols;
  }
  static std::string consensus_left(const std::vector<std::string>& seqs) {     // :100-108
    const std::string& s0 = seqs[0];
    if (seqs.size() == 1) return s0;
    size_t m = SIZE_MAX;
    for (const std::string& s : seqs) m = std::min(m, s.size());
    for (size_t i = 0; i < m; i++)
      for (size_t k = 1; k < seqs.size(); k++)
        if (seqs[k][i] != s0[i]) return s0.substr(0, i);
    return s0.substr(0, m);
  }
  static std::string consensus_right(const std::vector<std::string>& seqs) {    // :110-121
    const std::string& s0 = seqs[0];
    if (seqs.size() == 1) return s0;
    size_t m = SIZE_MAX;
    for (const std::string& s : seqs) m = std::min(m, s.size());
    for (size_t i = 0; i < m; i++)
      for (size_t k = 1; k < seqs.size(); k++) {
        const std::string& s = seqs[k];
        if (s[s.size() - 1 - i] != s0[s0.size() - 1 - i]) return s0.substr(s0.size() - i);
      }
    return s0.substr(s0.size() - m);
  }
  bool partial_matches_reference(int32_t start, const std::string& fl, const std::string& fr, int32_t stop) {  // :255-262
    if ((int64_t)start + (int64_t)fl.size() > R->stop || (int64_t)stop - (int64_t)fr.size() < R->start) return false;
    const std::string a = R->sub_string(start, (int32_t)fl.size());
    const std::string b = R->sub_string(stop - (int32_t)fr.size(), (int32_t)fr.size());
    return fl == a && fr == b;
  }
  // assembleAcrossBreak (:49-98) with assembleIntoBreak (:123-158)
  AsmFix across(int32_t bstart, int32_t bstop, bool is_gap, std::string* loop_out) {
    const int32_t start = std::max(bstart - o.break_radius, R->start);
    const int32_t stop = (int32_t)std::min<int64_t>((int64_t)bstop + 1 + o.break_radius, R->stop);
    const std::string left = R->sub_string(start, bstart - start);
    const std::string right = R->sub_string(bstop + 1, stop - bstop - 1);
    if (R->edge) return AsmFix();
    const std::vector<std::string> fw = A->try_forward(left);
    std::vector<std::string> rv = A->try_forward(asm_rc(right));
    for (std::string& s : rv) s = asm_rc(s);
    const std::string loop = A->loop_sequence;
    *loop_out = loop;
    const std::vector<AsmFix> sols = break_joins(start, fw, rv, stop);
    if (R->edge) return AsmFix();
    AsmFix sol;
    if (sols.size() == 1 || (o.multi_closure && sols.size() > 1)) sol = sols[0];
    bool ok = !asm_none(sol) && (loop.empty() || !o.tr_safe);
    if (ok && is_gap) {
      const int64_t diff = std::llabs((int64_t)sol.patch.size() - ((int64_t)bstop + 1 - bstart));
      if (diff > o.gap_margin) ok = false;
    }
    if (ok) return sol;
    if (is_gap || (o.fix_breaks && loop.empty())) {
      const std::string fr = consensus_right(rv), fl = consensus_left(fw);
      const int64_t new_start = (int64_t)start + (int64_t)fl.size(), new_stop = (int64_t)stop - (int64_t)fr.size();
      const bool extended = new_start >= (int64_t)bstart + 20 || new_stop <= (int64_t)bstop - 20;   // minExtend
      if (extended && !partial_matches_reference(start, fl, fr, stop) && !R->edge) {
        const int64_t gap = is_gap ? std::max<int64_t>(o.min_gap, new_stop - new_start) : o.min_gap;
        AsmFix s = trim_patch(start, fl + std::string((size_t)gap, 'N') + fr, stop);
        if (!s.patch.empty()) return s;
      }
    }
    return AsmFix();
  }
};

}  // namespace

extern "C" {

void pl_asm_opts_default(pl_asm_opts* o) {
  if (!o) return;
  o->k = 47;
  o->min_depth = 5;
  o->break_radius = 3 * 47;
  o->tr_safe = 1;
  o->multi_closure = 0;
  o->fix_breaks = 0;
  o->gap_margin = 100000;
  o->min_gap = 10;
}

pl_status pl_assemble_break(pl_ctx* c, const pl_read_soa* reads, const char* region_bases, const char* contig,
                            int64_t contig_len, int32_t region_start, int32_t region_stop, int32_t break_start,
                            int32_t break_stop, int32_t is_gap, const pl_asm_opts* opts, pl_asm_fix* out) {
  if (!c || !reads || !region_bases || !contig || !out) return fail(PL_E_INVAL, "null argument");
  std::memset(out, 0, sizeof *out);
  pl_asm_opts o;
  if (opts) o = *opts; else pl_asm_opts_default(&o);
  if (o.k < 1 || o.k > 8 * PL_KMER_WORDS) return fail(PL_E_INVAL, "K = %d outside [1, %d]", o.k, 8 * PL_KMER_WORDS);
  if (region_start < 1 || region_stop < region_start || region_stop > contig_len)
    return fail(PL_E_INVAL, "region %d-%d outside a contig of %lld", region_start, region_stop, (long long)contig_len);
  if (break_start > break_stop) return fail(PL_E_INVAL, "break %d-%d", break_start, break_stop);
  char* km = nullptr;
  uint8_t *nx = nullptr, *al = nullptr;
  int64_t nk = 0;
  double ms = 0;
  pl_status s = kmer_graph(c, reads, o.k, o.min_depth, &km, &nx, &al, &nk, &ms);
  if (s) return s;
  Assembler A;
  A.K = o.k;
  A.tr_safe = o.tr_safe != 0;
  A.kg.reserve((size_t)nk * 2);
  for (int64_t i = 0; i < nk; i++) {                           // buildGraph (Assembler.scala:131-150)
    const std::string k(km + (size_t)o.k * i, (size_t)o.k);
    const std::string prefix = k.substr(1);
    A.kg[k] = A.intern(prefix + (char)nx[i]);
    if (al[i]) A.alt[k] = A.intern(prefix + (char)al[i]);
  }
  free(km); free(nx); free(al);
  AsmRegion R;
  R.start = region_start; R.stop = region_stop; R.bases = region_bases; R.contig = contig;
  GapFill g;
  g.R = &R; g.A = &A; g.o = o;
  std::string loop;
  AsmFix f = g.across(break_start, break_stop, is_gap != 0, &loop);
  if (R.edge) return fail(PL_E_EDGE, "Pilon 1.24 throws here: a substring outside region %d-%d", region_start, region_stop);
  auto dup = [](const std::string& x) { char* p = (char*)malloc(x.size() + 1); if (p) { std::memcpy(p, x.data(), x.size()); p[x.size()] = 0; } return p; };
  out->start = f.start;
  out->ref = dup(f.ref);
  out->patch = dup(f.patch);
  out->tandem_repeat = dup(loop);
  out->ref_len = (int64_t)f.ref.size();
  out->patch_len = (int64_t)f.patch.size();
  if (!out->ref || !out->patch || !out->tandem_repeat) {
    free(out->ref); free(out->patch); free(out->tandem_repeat);
    std::memset(out, 0, sizeof *out);
    return fail(PL_E_OOM, "host allocation failed");
  }
  return PL_OK;
}

}  // extern "C"
