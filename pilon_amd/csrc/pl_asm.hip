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

// One fillGap / fixBreak given the recruited reads: the graph on the device,
// the walk and joins on the host.  PL_E_EDGE where Pilon throws.
pl_status assemble_core(pl_ctx* c, const pl_read_soa* reads, const char* region_bases, const char* contig,
                        int32_t region_start, int32_t region_stop, int32_t break_start, int32_t break_stop, bool is_gap,
                        const pl_asm_opts& o, AsmFix* fix, std::string* loop) {
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
  *fix = g.across(break_start, break_stop, is_gap, loop);
  if (R.edge) return fail(PL_E_EDGE, "Pilon 1.24 throws here: a substring outside region %d-%d", region_start, region_stop);
  return PL_OK;
}

pl_status check_opts(const pl_asm_opts& o) {
  if (o.k < 1 || o.k > 8 * PL_KMER_WORDS) return fail(PL_E_INVAL, "K = %d outside [1, %d]", o.k, 8 * PL_KMER_WORDS);
  if (o.break_radius < 0 || o.gap_margin < 0 || o.min_gap < 0) return fail(PL_E_INVAL, "negative assembly option");
  return PL_OK;
}

// Reads of several batches gathered into one SoA: the reads recruited for a
// break.
struct SoaGather {
  std::vector<int32_t> pos, tlen;
  std::vector<uint16_t> flag;
  std::vector<uint8_t> mapq;
  std::vector<int64_t> cigar_off{0}, seq_off{0}, exo_read, exo_off;
  std::vector<uint32_t> cigar;
  std::vector<uint8_t> bases2, nib;
  std::vector<int8_t> quals;
  int64_t n_nib = 0;

  void add(const pl_read_soa& S, int64_t i) {
    const int64_t s0 = S.seq_off[i], L = S.seq_off[i + 1] - s0, d0 = seq_off.back();
    const int64_t r = (int64_t)pos.size();
    pos.push_back(S.pos[i]);
    tlen.push_back(S.tlen ? S.tlen[i] : 0);
    flag.push_back(S.flag[i]);
    mapq.push_back(S.mapq[i]);
    if (S.cigar_off) {
      for (int64_t k = S.cigar_off[i]; k < S.cigar_off[i + 1]; k++) cigar.push_back(S.cigar[k]);
    }
    cigar_off.push_back((int64_t)cigar.size());
    seq_off.push_back(d0 + L);
    bases2.resize((size_t)((d0 + L + 3) / 4), 0);
    for (int64_t j = 0; j < L; j++) {
      const int64_t a = s0 + j, b = d0 + j;
      const uint8_t v = (S.bases2[a >> 2] >> (2 * (a & 3))) & 3u;
      bases2[(size_t)(b >> 2)] |= (uint8_t)(v << (2 * (b & 3)));
    }
    quals.insert(quals.end(), S.quals + s0, S.quals + s0 + L);
    // a read with non-ACGT codes: its nibbles too (exo_read is ascending)
    if (S.n_exotic > 0) {
      const int64_t* e = std::lower_bound(S.exo_read, S.exo_read + S.n_exotic, i);
      if (e != S.exo_read + S.n_exotic && *e == i) {
        const int64_t o0 = S.exo_off[e - S.exo_read];
        exo_read.push_back(r);
        exo_off.push_back(n_nib);
        nib.resize((size_t)((n_nib + L + 1) / 2), 0);
        for (int64_t j = 0; j < L; j++) {
          const int64_t a = o0 + j, b = n_nib + j;
          const uint8_t v = (S.nib[a >> 1] >> ((a & 1) ? 0 : 4)) & 15u;
          nib[(size_t)(b >> 1)] |= (uint8_t)(v << ((b & 1) ? 0 : 4));
        }
        n_nib += L;
      }
    }
  }
  pl_read_soa soa() const {
    pl_read_soa o;
    std::memset(&o, 0, sizeof o);
    o.n_reads = (int64_t)pos.size();
    o.pos = pos.data(); o.tlen = tlen.data(); o.flag = flag.data(); o.mapq = mapq.data();
    o.cigar_off = cigar_off.data(); o.cigar = cigar.data(); o.seq_off = seq_off.data();
    o.bases2 = bases2.data(); o.quals = quals.data();
    o.n_exotic = (int64_t)exo_read.size();
    o.exo_read = exo_read.data(); o.exo_off = exo_off.data(); o.nib = nib.data();
    return o;
  }
};

// BamFile.validateRead (BamFile.scala:100-104)
bool valid_read(const pl_config& cf, uint16_t f) {
  return (cf.nonpf || !(f & 0x200)) && (cf.duplicates || !(f & 0x400)) && !(f & 0x100);
}

// BamFile.readsInRegion over flankRegion (BamFile.scala:339-364): the valid
// reads whose span [pos, pos + reference length - 1] (one base for an
// unmapped or zero-length alignment) overlaps [max(start - flank, 1),
// stop + flank]
void recruit_flank(const pl_config& cf, const pl_read_soa& S, int32_t start, int32_t stop, int32_t flank, SoaGather& out) {
  const int64_t fs = std::max<int64_t>((int64_t)start - flank, 1), fe = (int64_t)stop + flank;
  for (int64_t i = 0; i < S.n_reads; i++) {
    if (!valid_read(cf, S.flag[i])) continue;
    const int64_t ps = S.pos[i];
    if (ps > fe) continue;
    int64_t rl = 0;
    if (S.cigar_off)
      for (int64_t k = S.cigar_off[i]; k < S.cigar_off[i + 1]; k++) {
        const uint32_t op = S.cigar[k] & 15u;
        if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rl += S.cigar[k] >> 4;
      }
    const int64_t pe = ((S.flag[i] & 0x4) || rl == 0) ? ps : ps + rl - 1;
    if (pe < fs) continue;
    out.add(S, i);
  }
}

// GenomeRegion.logFix (GenomeRegion.scala:472-502): the log line and the
// reassemblyFixes outcome of one attempt
plh::Outcome log_fix(const std::string& name, int32_t rs, int32_t re, const AsmFix& f, int32_t gap_size,
                     const std::string& loop, bool hash_prefix) {
  const int64_t n_ref = std::count(f.ref.begin(), f.ref.end(), 'N');
  const int64_t n_patch = std::count(f.patch.begin(), f.patch.end(), 'N');
  std::string msg;
  if (f.start == 0) msg = "NoSolution";
  else if (gap_size == 0 && f.ref.empty() && f.patch.empty()) msg = "NoChange";
  else if (gap_size == 0 && n_patch == 0) msg = "BreakFix";
  else if (n_patch > 0 && n_ref == 0) msg = "OpenedGap";
  else if (gap_size > 0 && n_ref == gap_size && n_patch == 0) msg = "ClosedGap";
  else if (gap_size > 0) msg = "PartialFill";
  else msg = "Unknown!";
  plh::Outcome o;
  o.start = rs;
  o.stop = re;
  o.outcome = loop.empty() ? msg : std::string("TandemRepeat");
  if (!loop.empty()) msg += " TandemRepeat " + std::to_string(loop.size());
  o.line = std::string(hash_prefix ? "# " : "") + "fix " + (gap_size > 0 ? "gap" : "break") + ": " +
           plh::region_string(name, rs, re) + " " + std::to_string(f.start) + " -" +
           std::to_string((int64_t)f.ref.size() - n_ref) + " +" + std::to_string((int64_t)f.patch.size() - n_patch) +
           " " + msg + "\n";
  return o;
}

// The issue regions of a finished region (gaps, possible breaks).
pl_status region_issue_lists(pl_region* r, std::vector<std::pair<int32_t, int32_t>> lists[7]) {
  return on_region_lane(r, [&](pl_ctx* c, Lane* L, const std::vector<std::string>&, const std::vector<int32_t>& contigs) {
    VcfBufs B;
    TrackArgs T;
    double ms = 0;
    if (pl_status s = tracks_of_lane(c, L, contigs, B, T, &ms, true)) return s;
    std::vector<uint8_t> mask((size_t)L->G);
    HIPCHK(hipMemcpyAsync(mask.data(), T.mask, (size_t)L->G, hipMemcpyDeviceToHost, L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
    const JobDev& J = L->hjobs[0];
    issue_lists(mask.data() + J.out, J.start, J.stop, lists);
    return PL_OK;
  });
}

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
  o->fix_gaps = 1;
  o->fix_local = 1;
}

pl_status pl_assemble_break(pl_ctx* c, const pl_read_soa* reads, const char* region_bases, const char* contig,
                            int64_t contig_len, int32_t region_start, int32_t region_stop, int32_t break_start,
                            int32_t break_stop, int32_t is_gap, const pl_asm_opts* opts, pl_asm_fix* out) {
  if (!c || !reads || !region_bases || !contig || !out) return fail(PL_E_INVAL, "null argument");
  std::memset(out, 0, sizeof *out);
  pl_asm_opts o;
  if (opts) o = *opts; else pl_asm_opts_default(&o);
  if (pl_status s = check_opts(o)) return s;
  if (region_start < 1 || region_stop < region_start || region_stop > contig_len)
    return fail(PL_E_INVAL, "region %d-%d outside a contig of %lld", region_start, region_stop, (long long)contig_len);
  if (break_start > break_stop) return fail(PL_E_INVAL, "break %d-%d", break_start, break_stop);
  AsmFix f;
  std::string loop;
  if (pl_status s = assemble_core(c, reads, region_bases, contig, region_start, region_stop, break_start, break_stop,
                                  is_gap != 0, o, &f, &loop))
    return s;
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

pl_status pl_region_reassemble(pl_region* r, const pl_read_soa* const* sets, const int32_t* set_types,
                               const int32_t* max_insert, int32_t n_sets, const pl_asm_opts* opts, int32_t* n_fixes) {
  if (!r || (n_sets > 0 && (!sets || !set_types || !max_insert)) || n_sets < 0) return fail(PL_E_INVAL, "null argument");
  if (!r->finished) return fail(PL_E_INVAL, "region not finished (pl_region_finish first)");
  pl_asm_opts o;
  if (opts) o = *opts; else pl_asm_opts_default(&o);
  if (pl_status s = check_opts(o)) return s;
  pl_ctx* c = r->ctx;
  const pl_config& cf = c->pc;
  for (int32_t k = 0; k < n_sets; k++) {
    if (!sets[k]) return fail(PL_E_INVAL, "read set %d is null", k);
    if (set_types[k] < PL_BAM_FRAGS || set_types[k] > PL_BAM_PACBIO) return fail(PL_E_INVAL, "read set %d: BAM type %d", k, set_types[k]);
    if (set_types[k] == PL_BAM_JUMPS)
      return fail(PL_E_INVAL, "read set %d is a jumps BAM: recruitBadMates pairs reads by name, which a read set does not carry", k);
    if (max_insert[k] < 0) return fail(PL_E_INVAL, "read set %d: negative maxInsertSize", k);
  }
  r->big.clear();
  r->outcomes.clear();
  r->vcf_embed.clear();
  std::vector<std::pair<int32_t, int32_t>> lists[7];
  if (pl_status s = region_issue_lists(r, lists)) return s;
  const Contig& ct = c->contigs[r->contig];
  const std::string rb((const char*)ct.host.data() + r->start - 1, (size_t)(r->stop - r->start + 1));
  const std::string bases = plh::snp_fixed_bases(cf, r->start, rb, r->changes);   // fixIssues(snpFixList)
  const char* contig = (const char*)ct.host.data();
  // recruitReads (GapFiller.scala:364-403, --nostrays): each frags BAM's
  // flank reads, then each unpaired BAM's unless a long-read run
  auto recruit = [&](int32_t bs, int32_t be, SoaGather& g) {
    for (int32_t k = 0; k < n_sets; k++)
      if (set_types[k] == PL_BAM_FRAGS) recruit_flank(cf, *sets[k], bs, be, max_insert[k], g);
    if (!cf.longread)
      for (int32_t k = 0; k < n_sets; k++)
        if (set_types[k] != PL_BAM_FRAGS) recruit_flank(cf, *sets[k], bs, be, max_insert[k], g);
  };
  // fillGap over the gaps, then fixBreak over the possible breaks
  // (GenomeRegion.scala:382-410)
  for (int pass = 0; pass < 2; pass++) {
    const bool is_gap = pass == 0;
    if (is_gap ? !o.fix_gaps : !o.fix_local) continue;
    for (const auto& br : lists[is_gap ? ISSUE_GAPS : ISSUE_BREAKS]) {
      SoaGather g;
      recruit(br.first, br.second, g);
      const pl_read_soa soa = g.soa();
      AsmFix f;
      std::string loop;
      if (pl_status s = assemble_core(c, &soa, bases.data(), contig, r->start, r->stop, br.first, br.second, is_gap, o,
                                      &f, &loop)) {
        r->big.clear();
        r->outcomes.clear();
        return s;
      }
      const int32_t gap_size = is_gap ? br.second + 1 - br.first : 0;
      if (is_gap) {
        if (f.start > 0) {
          r->big.push_back({f.start, f.ref, f.patch});
          r->outcomes.push_back(log_fix(ct.name, br.first, br.second, f, gap_size, loop, false));
        }
      } else if (f.start > 0 && std::max(f.ref.size(), f.patch.size()) > 10) {
        r->big.push_back({f.start, f.ref, f.patch});
        r->outcomes.push_back(log_fix(ct.name, br.first, br.second, f, 0, loop, false));
      } else if (f.start == 0) {
        r->outcomes.push_back(log_fix(ct.name, br.first, br.second, f, 0, loop, true));
      }
    }
  }
  if (n_fixes) *n_fixes = (int32_t)r->big.size();
  return PL_OK;
}

pl_status pl_region_reassembly_log(pl_region* r, char** text, int64_t* len) {
  if (!r || !text || !len) return fail(PL_E_INVAL, "null argument");
  std::string t;
  for (const plh::Outcome& o : r->outcomes) t += o.line;
  *text = (char*)malloc(t.size() + 1);
  if (!*text) return fail(PL_E_OOM, "host allocation failed");
  std::memcpy(*text, t.c_str(), t.size() + 1);
  *len = (int64_t)t.size();
  return PL_OK;
}

}  // extern "C"
