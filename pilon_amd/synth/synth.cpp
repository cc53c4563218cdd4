// synth.cpp — seeded synthetic drafts, truths and aligned reads for the
// pileup benchmarks (the workloads of the five benchmark configurations).
//
// A draft contig is random ACGT at a chosen GC fraction (optionally with
// repeat copies, hotspot windows and heterozygous sites).  The truth is the
// draft with planted SNVs and indels.  Read pairs are drawn from the truth and
// carry their TRUE alignment to the draft as an M/I/D/S CIGAR, exactly as an
// ideal aligner would report it; errors (substitutions, 1..k bp indels, soft
// clips, improper pairs) are added on top.  Output is the packed
// structure-of-arrays of include/pilon_hip.h, coordinate-sorted per contig.
//
// Host-only C++ (no GPU code); built into libpilon_synth.so.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>
#include <omp.h>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <zlib.h>
#include <cstdio>
#include <string>

#include "../../include/pilon_hip.h"

namespace {

struct Rng {  // xoshiro256** seeded through splitmix64
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (int i = 0; i < 4; i++) {
      seed += 0x9e3779b97f4a7c15ULL;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
    return r;
  }
  double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  int64_t range(int64_t lo, int64_t hi) {  // inclusive
    return lo + (int64_t)(uni() * (double)(hi - lo + 1));
  }
  double normal() {
    double u1 = uni(), u2 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  double gamma(double k) {  // Marsaglia-Tsang, k >= 1
    double d = k - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x = normal(), v = 1.0 + c * x;
      if (v <= 0) continue;
      v = v * v * v;
      double u = uni();
      if (u < 1 - 0.0331 * x * x * x * x) return d * v;
      if (std::log(u) < 0.5 * x * x + d * (1 - v + std::log(v))) return d * v;
    }
  }
  double beta(double a, double b) { double x = gamma(a), y = gamma(b); return x / (x + y); }
};

uint64_t mix(uint64_t a, uint64_t b) {
  uint64_t z = a * 0x9e3779b97f4a7c15ULL + b + 0x632be59bd9b4e019ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

const char ACGT[4] = {'A', 'C', 'G', 'T'};

uint8_t rand_base(Rng& r, double gc) {
  double u = r.uni();
  if (u < gc) return u < gc / 2 ? 'C' : 'G';
  return (u - gc) < (1 - gc) / 2 ? 'A' : 'T';
}
uint8_t other_base(Rng& r, uint8_t b) {
  for (;;) { uint8_t c = ACGT[r.next() & 3]; if (c != b) return c; }
}
int code2(uint8_t c) { return c == 'A' ? 0 : c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : -1; }
int nib_code(uint8_t c) {
  static const char* N = "=ACMGRSVTWYHKDBN";
  for (int i = 0; i < 16; i++) if (N[i] == c) return i;
  return 15;
}

}  // namespace

extern "C" {

typedef struct synth_params {
  uint64_t seed;
  int32_t n_contigs;
  const int64_t* contig_lens;
  double gc;
  int32_t read_len;
  double coverage;           // per-base depth of reads (both mates)
  double insert_mean, insert_sd;
  double sub_rate;           // per-base substitution error
  double err_indel_rate;     // per-base error indel rate
  int32_t err_indel_max;
  double clip_frac;          // fraction of reads with an extra soft clip
  int32_t clip_min, clip_max;
  double improper_frac;      // fraction of pairs with a distant, improper mate
  int32_t n_snv, n_indel;    // planted per contig when >= 0
  double snv_rate, indel_rate;  // planted per base otherwise
  int32_t indel_max;
  int32_t repeat_copies, repeat_len;   // copies of one repeat per contig; reads on them get MAPQ 0
  int32_t qual_lo, qual_hi, qual_decay;
  int32_t mapq;
  double n_rate;             // per-base rate of 'N' in reads (exotic path)
  double het_rate;           // heterozygous SNV sites per base, AF ~ Beta(het_a, het_b)
  double het_a, het_b;
  int32_t n_hotspots, hotspot_len;  // windows with extra coverage
  double hotspot_cov;
  int32_t n_breakpoints;     // planted breakpoints; reads crossing them are clipped
  double bp_clip_frac;
  int32_t bp_clip_min, bp_clip_max;
  int32_t threads;
} synth_params;

}  // extern "C"

namespace {

struct Variant { int64_t pos; int32_t kind; int32_t len; };  // kind 0 snv,1 ins,2 del (draft coords)

struct RawRead {
  int32_t pos;      // 1-based alignment start
  int32_t aend;     // 1-based alignment end
  int32_t tlen;
  uint16_t flag;
  uint8_t mapq;
  uint32_t ncig, cig0;     // into chunk cigar
  uint64_t nseq, seq0;     // into chunk bases/quals
  bool exotic;
};

struct Chunk {
  std::vector<RawRead> reads;
  std::vector<uint32_t> cigar;
  std::vector<uint8_t> bases;   // ASCII
  std::vector<int8_t> quals;
};

struct ContigOut {
  std::vector<uint8_t> draft, truth;
  std::vector<Variant> variants;
  // packed reads
  std::vector<int32_t> pos, tlen;
  std::vector<uint16_t> flag;
  std::vector<uint8_t> mapq;
  std::vector<int64_t> cigar_off;
  std::vector<uint32_t> cigar;
  std::vector<int64_t> seq_off;
  std::vector<uint8_t> bases2;
  std::vector<int8_t> quals;
  std::vector<int64_t> exo_read, exo_off;
  std::vector<uint8_t> nib;
};

struct Het { int64_t d; uint8_t alt; float af; };

struct ContigModel {
  const synth_params* P;
  std::vector<uint8_t> draft, truth;
  std::vector<int64_t> t2d;          // truth index -> draft index, -1 if inserted
  std::vector<uint8_t> repeat_mask;  // per draft base
  std::vector<int64_t> het_pos;      // sorted draft indices
  std::vector<Het> hets;
  std::vector<int64_t> bp;           // sorted truth breakpoints
  std::vector<std::pair<int64_t, int64_t>> hot;  // truth intervals [a, b)
};

void add_op(std::vector<uint32_t>& cig, size_t c0, int op, uint32_t len) {
  if (len == 0) return;
  if (cig.size() > c0 && (cig.back() & 0xf) == (uint32_t)op) cig.back() += len << 4;
  else cig.push_back((len << 4) | (uint32_t)op);
}

// One read [t0, t0+L) on the truth, forward orientation; returns false if it
// aligns nowhere.  clip_lead/clip_trail: extra soft-clipped random bases.
bool make_read(const ContigModel& M, Rng& r, int64_t t0, int32_t L, bool reverse, bool mapq0_ok,
               int32_t clip_lead, int32_t clip_trail, Chunk& ch, RawRead& rr) {
  const synth_params& P = *M.P;
  const int64_t T = (int64_t)M.truth.size();
  // events: read base + draft index (-1 inserted, -2 soft clip)
  static thread_local std::vector<uint8_t> tl_eb;
  static thread_local std::vector<int64_t> tl_ed;
  std::vector<uint8_t>& eb = tl_eb;    // one thread-local lookup per read, not per base
  std::vector<int64_t>& ed = tl_ed;
  eb.clear(); ed.clear();
  for (int i = 0; i < clip_lead; i++) { eb.push_back(rand_base(r, 0.5)); ed.push_back(-2); }
  int64_t t = t0;
  int32_t body = L - clip_lead - clip_trail;
  // heterozygous sites at or after the read's draft positions, found once:
  // draft indices only grow along the read
  size_t hc = 0;
  bool hc_set = false;
  while ((int32_t)eb.size() < clip_lead + body && t < T) {
    if (P.err_indel_rate > 0 && r.uni() < P.err_indel_rate) {
      int k = (int)r.range(1, P.err_indel_max);
      if (r.next() & 1) {  // error insertion
        for (int i = 0; i < k && (int32_t)eb.size() < clip_lead + body; i++) { eb.push_back(rand_base(r, 0.5)); ed.push_back(-1); }
      } else {
        t += k;          // error deletion
      }
      continue;
    }
    uint8_t b = M.truth[t];
    int64_t d = M.t2d[t];
    if (d >= 0 && !M.het_pos.empty()) {
      if (!hc_set) {
        hc = (size_t)(std::lower_bound(M.het_pos.begin(), M.het_pos.end(), d) - M.het_pos.begin());
        hc_set = true;
      }
      while (hc < M.het_pos.size() && M.het_pos[hc] < d) hc++;
      if (hc < M.het_pos.size() && M.het_pos[hc] == d) {
        const Het& h = M.hets[hc];
        if (r.uni() < h.af) b = h.alt;
      }
    }
    if (P.sub_rate > 0 && r.uni() < P.sub_rate) b = other_base(r, b);
    if (P.n_rate > 0 && r.uni() < P.n_rate) b = 'N';
    eb.push_back(b); ed.push_back(d);
    t++;
  }
  for (int i = 0; i < clip_trail; i++) { eb.push_back(rand_base(r, 0.5)); ed.push_back(-2); }
  int32_t n = (int32_t)eb.size();
  // first/last aligned base
  int32_t f = -1, l = -1;
  for (int32_t i = 0; i < n; i++) if (ed[i] >= 0) { if (f < 0) f = i; l = i; }
  if (f < 0) return false;
  rr.cig0 = (uint32_t)ch.cigar.size();
  std::vector<uint32_t>& cig = ch.cigar;
  size_t c0 = cig.size();
  add_op(cig, c0, 4, (uint32_t)f);  // everything before the first aligned base is clipped
  int64_t prev = -1;
  bool mapq0 = false;
  for (int32_t i = f; i <= l; i++) {
    int64_t d = ed[i];
    if (d == -1 || d == -2) { add_op(cig, c0, 1, 1); continue; }  // a clip base inside becomes I
    if (prev >= 0 && d > prev + 1) add_op(cig, c0, 2, (uint32_t)(d - prev - 1));
    add_op(cig, c0, 0, 1);
    if (mapq0_ok && !M.repeat_mask.empty() && M.repeat_mask[d]) mapq0 = true;
    prev = d;
  }
  add_op(cig, c0, 4, (uint32_t)(n - 1 - l));  // and everything after the last one
  rr.ncig = (uint32_t)(cig.size() - c0);
  rr.pos = (int32_t)(ed[f] + 1);
  rr.aend = (int32_t)(prev + 1);
  rr.mapq = mapq0 ? 0 : (uint8_t)P.mapq;
  rr.seq0 = ch.bases.size();
  rr.nseq = (uint64_t)n;
  rr.exotic = false;
  for (int32_t i = 0; i < n; i++) {
    ch.bases.push_back(eb[i]);
    if (code2(eb[i]) < 0) rr.exotic = true;
  }
  // qualities in read orientation: U{lo..hi} - floor(decay * p), p toward 3'
  for (int32_t i = 0; i < n; i++) {
    int32_t ri = reverse ? (n - 1 - i) : i;
    double p = n > 1 ? (double)ri / (double)(n - 1) : 0.0;
    int32_t q = (int32_t)r.range(P.qual_lo, P.qual_hi) - (int32_t)std::floor(P.qual_decay * p);
    if (q < 2) q = 2;
    ch.quals.push_back((int8_t)q);
  }
  return true;
}

void build_contig(const synth_params& P, int ci, ContigModel& M, std::vector<Variant>& vars) {
  Rng r(mix(P.seed, 0x1000 + (uint64_t)ci));
  int64_t Ld = P.contig_lens[ci];
  M.draft.resize((size_t)Ld);
  for (int64_t i = 0; i < Ld; i++) M.draft[i] = rand_base(r, P.gc);
  if (P.repeat_copies > 0 && P.repeat_len > 0 && Ld > (int64_t)P.repeat_len * P.repeat_copies * 3) {
    std::vector<uint8_t> rep((size_t)P.repeat_len);
    for (auto& b : rep) b = rand_base(r, P.gc);
    M.repeat_mask.assign((size_t)Ld, 0);
    int64_t slot = Ld / P.repeat_copies;
    for (int k = 0; k < P.repeat_copies; k++) {
      int64_t at = k * slot + r.range(0, slot - P.repeat_len - 1);
      std::memcpy(&M.draft[at], rep.data(), rep.size());
      std::memset(&M.repeat_mask[at], 1, rep.size());
    }
  }
  // planted variants, kept 30+ bp apart and away from the ends
  int64_t nsnv = P.n_snv >= 0 ? P.n_snv : (int64_t)std::llround(P.snv_rate * Ld);
  int64_t nind = P.n_indel >= 0 ? P.n_indel : (int64_t)std::llround(P.indel_rate * Ld);
  std::vector<int64_t> spots;
  int64_t want = Ld >= 1000 ? nsnv + nind : 0, tries = 0;
  std::vector<uint8_t> used((size_t)Ld, 0);
  while ((int64_t)spots.size() < want && tries < want * 50 + 100) {
    tries++;
    int64_t p = r.range(200, Ld - 200 - 1);
    if (p < 200 || p >= Ld - 200) break;
    bool ok = true;
    for (int64_t q = std::max<int64_t>(0, p - 30); q <= std::min<int64_t>(Ld - 1, p + 30); q++) if (used[q]) { ok = false; break; }
    if (!ok) continue;
    used[p] = 1;
    spots.push_back(p);
  }
  for (size_t k = 0; k < spots.size(); k++) {
    Variant v;
    v.pos = spots[k];
    if ((int64_t)k < nsnv) { v.kind = 0; v.len = 1; }
    else { v.kind = (r.next() & 1) ? 1 : 2; v.len = (int32_t)r.range(1, P.indel_max); }
    vars.push_back(v);
  }
  std::sort(vars.begin(), vars.end(), [](const Variant& a, const Variant& b) { return a.pos < b.pos; });
  // truth + truth->draft map
  M.truth.reserve((size_t)Ld + 1024);
  M.t2d.reserve((size_t)Ld + 1024);
  size_t vi = 0;
  for (int64_t d = 0; d < Ld;) {
    if (vi < vars.size() && vars[vi].pos == d) {
      const Variant& v = vars[vi++];
      if (v.kind == 0) {
        M.truth.push_back(other_base(r, M.draft[d])); M.t2d.push_back(d); d++;
      } else if (v.kind == 1) {       // insertion after draft base d
        M.truth.push_back(M.draft[d]); M.t2d.push_back(d);
        for (int i = 0; i < v.len; i++) { M.truth.push_back(rand_base(r, P.gc)); M.t2d.push_back(-1); }
        d++;
      } else {                         // deletion of draft bases d+1 .. d+len
        M.truth.push_back(M.draft[d]); M.t2d.push_back(d);
        d += 1 + v.len;
      }
      continue;
    }
    M.truth.push_back(M.draft[d]); M.t2d.push_back(d); d++;
  }
  if (P.het_rate > 0) {
    int64_t nh = (int64_t)std::llround(P.het_rate * Ld);
    for (int64_t k = 0; k < nh; k++) {
      int64_t d = r.range(100, Ld - 101);
      M.hets.push_back({d, other_base(r, M.draft[d]), (float)r.beta(P.het_a, P.het_b)});
    }
    std::sort(M.hets.begin(), M.hets.end(), [](const Het& a, const Het& b) { return a.d < b.d; });
    M.hets.erase(std::unique(M.hets.begin(), M.hets.end(), [](const Het& a, const Het& b) { return a.d == b.d; }), M.hets.end());
    for (auto& h : M.hets) M.het_pos.push_back(h.d);
  }
  int64_t T = (int64_t)M.truth.size();
  for (int k = 0; k < P.n_hotspots && T > P.hotspot_len + 1000; k++) {
    int64_t a = r.range(500, T - P.hotspot_len - 500);
    M.hot.push_back({a, a + P.hotspot_len});
  }
  for (int k = 0; k < P.n_breakpoints && T > 2000; k++) M.bp.push_back(r.range(500, T - 500));
  std::sort(M.bp.begin(), M.bp.end());
  M.P = &P;
}

// clip lengths for a read on [t0, t0+L): planted breakpoints inside the read
// clip the part beyond the breakpoint with probability bp_clip_frac.
void pick_clips(const ContigModel& M, Rng& r, int64_t t0, int32_t L, int32_t& lead, int32_t& trail) {
  const synth_params& P = *M.P;
  lead = trail = 0;
  if (P.clip_frac > 0 && r.uni() < P.clip_frac) {
    int32_t k = (int32_t)r.range(P.clip_min, P.clip_max);
    if (k > L / 2) k = L / 2;
    if (r.next() & 1) lead = k; else trail = k;
  }
  if (!M.bp.empty() && P.bp_clip_frac > 0) {
    auto it = std::lower_bound(M.bp.begin(), M.bp.end(), t0);
    if (it != M.bp.end() && *it < t0 + L && r.uni() < P.bp_clip_frac) {
      int32_t k = (int32_t)r.range(P.bp_clip_min, P.bp_clip_max);
      if (k > L - 20) k = L - 20;
      if (r.next() & 1) lead = std::max(lead, k); else trail = std::max(trail, k);
    }
  }
}

void gen_pairs(const ContigModel& M, int ci, int64_t chunk_id, int64_t npairs, int64_t lo, int64_t hi,
               double /*cov*/, Chunk& ch) {
  const synth_params& P = *M.P;
  Rng r(mix(mix(P.seed, 0x2000 + (uint64_t)ci), (uint64_t)chunk_id));
  const int64_t T = (int64_t)M.truth.size();
  const int32_t L = P.read_len;
  ch.reads.reserve((size_t)(2 * npairs));
  ch.bases.reserve((size_t)(2 * npairs * (L + 4)));
  ch.quals.reserve((size_t)(2 * npairs * (L + 4)));
  ch.cigar.reserve((size_t)(2 * npairs * 3));
  for (int64_t k = 0; k < npairs; k++) {
    int64_t F = (int64_t)std::llround(P.insert_mean + P.insert_sd * r.normal());
    if (F < L) F = L;
    if (F > T) F = T;
    int64_t a = std::max<int64_t>(0, lo), b = std::min<int64_t>(T - F, hi - 1);
    if (b < a) continue;
    int64_t t0 = r.range(a, b);
    bool improper = P.improper_frac > 0 && r.uni() < P.improper_frac;
    bool firstIsFwd = (r.next() & 1) != 0;
    RawRead r1{}, r2{};
    int32_t cl, ct;
    pick_clips(M, r, t0, L, cl, ct);
    bool ok1 = make_read(M, r, t0, L, false, true, cl, ct, ch, r1);
    int64_t t2 = t0 + F - L;
    bool rev2 = true;
    if (improper) {
      t2 = r.range(0, T - L);
      rev2 = (r.next() & 1) != 0;
    }
    pick_clips(M, r, t2, L, cl, ct);
    bool ok2 = make_read(M, r, t2, L, rev2, true, cl, ct, ch, r2);
    if (!ok1 || !ok2) continue;  // drop the pair; unreferenced chunk bytes are harmless
    uint16_t common = 0x1 | (improper ? 0 : 0x2) | PL_RF_MATE_SAME_REF;
    r1.flag = common | (rev2 ? 0x20 : 0) | (firstIsFwd ? 0x40 : 0x80);
    r2.flag = common | (rev2 ? 0x10 : 0) | (firstIsFwd ? 0x80 : 0x40);
    if (improper) { r1.tlen = 0; r2.tlen = 0; }
    else {
      int32_t ls = std::min(r1.pos, r2.pos), re = std::max(r1.aend, r2.aend);
      int32_t isz = re - ls + 1;
      bool r1left = r1.pos < r2.pos || (r1.pos == r2.pos && firstIsFwd);
      r1.tlen = r1left ? isz : -isz;
      r2.tlen = r1left ? -isz : isz;
    }
    ch.reads.push_back(r1);
    ch.reads.push_back(r2);
  }
}

}  // namespace

struct SynthResult {
  std::vector<ContigOut> contigs;
};

extern "C" {

void synth_default_params(synth_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->seed = 1001; p->gc = 0.5; p->read_len = 150; p->coverage = 40;
  p->insert_mean = 400; p->insert_sd = 40; p->sub_rate = 0.003;
  p->err_indel_max = 3; p->clip_min = 10; p->clip_max = 60;
  p->n_snv = -1; p->n_indel = -1; p->indel_max = 10;
  p->qual_lo = 20; p->qual_hi = 40; p->qual_decay = 10; p->mapq = 60;
  p->het_a = 20; p->het_b = 20; p->hotspot_len = 10000; p->bp_clip_min = 30; p->bp_clip_max = 200;
  p->threads = 0;
}

// PL_SYNTH_PROGRESS=1 in the environment: a line on stderr per phase and
// per tenth of the pair jobs (large workloads take minutes)
static bool synth_progress() {
  const char* e = getenv("PL_SYNTH_PROGRESS");
  return e && *e && *e != '0';
}

void* synth_generate(const synth_params* Pin) {
  synth_params P = *Pin;
  const bool prog = synth_progress();
  SynthResult* res = new SynthResult();
  res->contigs.resize((size_t)P.n_contigs);
  int nthreads = P.threads > 0 ? P.threads : omp_get_max_threads();
  std::vector<ContigModel> models((size_t)P.n_contigs);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int ci = 0; ci < P.n_contigs; ci++) {
    std::vector<Variant> vars;
    build_contig(P, ci, models[ci], vars);
    res->contigs[ci].variants = vars;
  }
  // fragment chunks: ~200k pairs each, deterministic seeds
  struct Job { int ci; int64_t chunk; int64_t npairs; int64_t lo, hi; };
  std::vector<Job> jobs;
  const int64_t PER = 100000;
  for (int ci = 0; ci < P.n_contigs; ci++) {
    int64_t T = (int64_t)models[ci].truth.size();
    int64_t npairs = (int64_t)std::llround(P.coverage * (double)T / (2.0 * P.read_len));
    int64_t nch = std::max<int64_t>(1, (npairs + PER - 1) / PER);
    for (int64_t c = 0; c < nch; c++) {
      int64_t lo = T * c / nch, hi = T * (c + 1) / nch;
      int64_t np = npairs * (c + 1) / nch - npairs * c / nch;
      jobs.push_back({ci, c, np, c == 0 ? 0 : lo, c == nch - 1 ? T : hi});
    }
    for (size_t h = 0; h < models[ci].hot.size(); h++) {
      int64_t a = models[ci].hot[h].first, b = models[ci].hot[h].second;
      int64_t extra = (int64_t)std::llround(P.hotspot_cov * (double)(b - a) / (2.0 * P.read_len));
      jobs.push_back({ci, 1000000 + (int64_t)h, extra, std::max<int64_t>(0, a - (int64_t)P.insert_mean), b});
    }
  }
  std::vector<Chunk> chunks(jobs.size());
  if (prog) fprintf(stderr, "synth: %d contigs built, %zu pair jobs\n", P.n_contigs, jobs.size());
  std::atomic<size_t> jobs_done{0};
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (size_t j = 0; j < jobs.size(); j++) {
    gen_pairs(models[jobs[j].ci], jobs[j].ci, jobs[j].chunk, jobs[j].npairs, jobs[j].lo, jobs[j].hi,
              P.coverage, chunks[j]);
    const size_t k = ++jobs_done;
    if (prog && (k * 10 / jobs.size() != (k - 1) * 10 / jobs.size()))
      fprintf(stderr, "synth: %zu of %zu pair jobs\n", k, jobs.size());
  }
  // per contig: sort by pos (stable in generation order) and pack; a batch
  // of many contigs packs one contig per thread, a few large contigs pack
  // each one with all threads
  std::vector<std::vector<size_t>> jobs_of((size_t)P.n_contigs);
  for (size_t j = 0; j < jobs.size(); j++) jobs_of[(size_t)jobs[j].ci].push_back(j);
  struct Key { int32_t pos; uint32_t chunk; uint32_t read; };
  std::vector<std::vector<Key>> order((size_t)P.n_contigs);
  // reads of each contig in generation order (chunk, then read), sorted
  // stably by position: one contig per thread
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int ci = 0; ci < P.n_contigs; ci++) {
    std::vector<Key>& idx = order[(size_t)ci];
    size_t total = 0;
    for (size_t j : jobs_of[(size_t)ci]) total += chunks[j].reads.size();
    idx.reserve(total);
    for (size_t j : jobs_of[(size_t)ci])
      for (size_t k = 0; k < chunks[j].reads.size(); k++) idx.push_back({chunks[j].reads[k].pos, (uint32_t)j, (uint32_t)k});
    std::stable_sort(idx.begin(), idx.end(), [](const Key& a, const Key& b) { return a.pos < b.pos; });
  }
  auto pack = [&](int ci, int inner) {
    ContigOut& out = res->contigs[ci];
    out.draft = models[ci].draft;
    out.truth = models[ci].truth;
    std::vector<Key> idx;
    idx.swap(order[(size_t)ci]);
    size_t n = idx.size();
    out.pos.resize(n); out.tlen.resize(n); out.flag.resize(n); out.mapq.resize(n);
    out.cigar_off.resize(n + 1); out.seq_off.resize(n + 1);
    out.cigar_off[0] = 0; out.seq_off[0] = 0;
    for (size_t i = 0; i < n; i++) {
      const RawRead& rr = chunks[idx[i].chunk].reads[idx[i].read];
      out.cigar_off[i + 1] = out.cigar_off[i] + rr.ncig;
      out.seq_off[i + 1] = out.seq_off[i] + (int64_t)rr.nseq;
    }
    out.cigar.resize((size_t)out.cigar_off[n]);
    out.quals.resize((size_t)out.seq_off[n] + 8, 0);
    out.bases2.assign((size_t)(out.seq_off[n] + 3) / 4 + 8, 0);
    uint8_t* b2 = out.bases2.data();
#pragma omp parallel for schedule(static) num_threads(inner) if (inner > 1)
    for (size_t i = 0; i < n; i++) {
      const Chunk& ch = chunks[idx[i].chunk];
      const RawRead& rr = ch.reads[idx[i].read];
      out.pos[i] = rr.pos; out.tlen[i] = rr.tlen; out.mapq[i] = rr.mapq;
      out.flag[i] = rr.flag | (rr.exotic ? PL_RF_EXOTIC : 0);
      std::memcpy(&out.cigar[(size_t)out.cigar_off[i]], &ch.cigar[rr.cig0], sizeof(uint32_t) * rr.ncig);
      const int64_t so = out.seq_off[i], se = so + (int64_t)rr.nseq;
      std::memcpy(&out.quals[(size_t)so], &ch.quals[rr.seq0], rr.nseq);
      if (rr.exotic) continue;
      // 2-bit codes, four to a byte; bytes shared with a neighbouring read are ORed atomically
      uint8_t acc = 0;
      for (int64_t q = so; q < se; q++) {
        acc |= (uint8_t)(code2(ch.bases[rr.seq0 + (uint64_t)(q - so)]) << (2 * (q & 3)));
        if ((q & 3) == 3 || q == se - 1) {
          const int64_t byte = q >> 2;
          if (byte * 4 >= so && byte * 4 + 3 < se) b2[byte] = acc;
          else __atomic_fetch_or(&b2[byte], acc, __ATOMIC_RELAXED);
          acc = 0;
        }
      }
    }
    // reads with non-ACGT codes: nibbles, each read from a byte boundary
    for (size_t i = 0; i < n; i++) {
      const Chunk& ch = chunks[idx[i].chunk];
      const RawRead& rr = ch.reads[idx[i].read];
      if (!rr.exotic) continue;
      out.exo_read.push_back((int64_t)i);
      out.exo_off.push_back((int64_t)out.nib.size() * 2);
      std::vector<uint8_t> codes(rr.nseq);
      for (uint64_t k = 0; k < rr.nseq; k++) codes[k] = (uint8_t)nib_code(ch.bases[rr.seq0 + k]);
      for (uint64_t k = 0; k < rr.nseq; k += 2) {
        uint8_t hi = codes[k], lo = k + 1 < rr.nseq ? codes[k + 1] : 0;
        out.nib.push_back((uint8_t)((hi << 4) | lo));
      }
    }
    if (out.nib.empty()) out.nib.push_back(0);
    for (size_t j : jobs_of[(size_t)ci]) { Chunk().reads.swap(chunks[j].reads); std::vector<uint32_t>().swap(chunks[j].cigar);
      std::vector<uint8_t>().swap(chunks[j].bases); std::vector<int8_t>().swap(chunks[j].quals); }
  };
  if (prog) fprintf(stderr, "synth: reads sorted, packing\n");
  if (P.n_contigs >= 4 * nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int ci = 0; ci < P.n_contigs; ci++) pack(ci, 1);
  } else {
    for (int ci = 0; ci < P.n_contigs; ci++) {
      pack(ci, nthreads);
      if (prog) fprintf(stderr, "synth: contig %d of %d packed\n", ci + 1, P.n_contigs);
    }
  }
  return res;
}

void synth_free(void* h) { delete (SynthResult*)h; }

int64_t synth_contig(void* h, int32_t ci, const uint8_t** draft, const uint8_t** truth, int64_t* truth_len) {
  ContigOut& c = ((SynthResult*)h)->contigs[(size_t)ci];
  *draft = c.draft.data();
  *truth = c.truth.data();
  *truth_len = (int64_t)c.truth.size();
  return (int64_t)c.draft.size();
}

int64_t synth_variants(void* h, int32_t ci, int64_t* pos, int32_t* kind, int32_t* len, int64_t cap) {
  ContigOut& c = ((SynthResult*)h)->contigs[(size_t)ci];
  int64_t n = (int64_t)c.variants.size();
  for (int64_t i = 0; i < n && i < cap; i++) { pos[i] = c.variants[i].pos; kind[i] = c.variants[i].kind; len[i] = c.variants[i].len; }
  return n;
}

void synth_reads(void* h, int32_t ci, pl_read_soa* s) {
  ContigOut& c = ((SynthResult*)h)->contigs[(size_t)ci];
  s->n_reads = (int64_t)c.pos.size();
  s->pos = c.pos.data(); s->tlen = c.tlen.data(); s->flag = c.flag.data(); s->mapq = c.mapq.data();
  s->cigar_off = c.cigar_off.data(); s->cigar = c.cigar.empty() ? nullptr : c.cigar.data();
  s->seq_off = c.seq_off.data(); s->bases2 = c.bases2.data(); s->quals = c.quals.data();
  s->n_exotic = (int64_t)c.exo_read.size();
  s->exo_read = c.exo_read.empty() ? nullptr : c.exo_read.data();
  s->exo_off = c.exo_off.empty() ? nullptr : c.exo_off.data();
  s->nib = c.nib.data();
}

}  // extern "C"

// ---- the workload as a BAM file (SAMv1 sections 4.1-4.2): one record per
// read, named r<index>, the mate on the same reference iff
// PL_RF_MATE_SAME_REF, the SAM flag bits only; BGZF blocks of 0xFF00 bytes,
// raw deflate at `level`, and the end-of-file marker.  Returns the size of
// the inflated BAM stream, or -1.
namespace {
void put32(std::vector<uint8_t>& v, uint32_t x) { for (int k = 0; k < 4; k++) v.push_back((uint8_t)(x >> (8 * k))); }
void put16(std::vector<uint8_t>& v, uint16_t x) { v.push_back((uint8_t)x); v.push_back((uint8_t)(x >> 8)); }
uint32_t reg2bin(int64_t beg, int64_t end) {
  --end;
  if (beg >> 14 == end >> 14) return (uint32_t)(((1 << 15) - 1) / 7 + (beg >> 14));
  if (beg >> 17 == end >> 17) return (uint32_t)(((1 << 12) - 1) / 7 + (beg >> 17));
  if (beg >> 20 == end >> 20) return (uint32_t)(((1 << 9) - 1) / 7 + (beg >> 20));
  if (beg >> 23 == end >> 23) return (uint32_t)(((1 << 6) - 1) / 7 + (beg >> 23));
  if (beg >> 26 == end >> 26) return (uint32_t)(((1 << 3) - 1) / 7 + (beg >> 26));
  return 0;
}
void contig_records(const ContigOut& c, int32_t ref, std::vector<uint8_t>& out) {
  static const uint8_t kNib[4] = {1, 2, 4, 8};
  const size_t n = c.pos.size();
  size_t e = 0;
  for (size_t i = 0; i < n; i++) {
    const int64_t s0 = c.seq_off[i], L = c.seq_off[i + 1] - s0;
    const int64_t c0 = c.cigar_off[i], nc = c.cigar_off[i + 1] - c0;
    const std::string name = "r" + std::to_string(i);
    int64_t rlen = 0;
    for (int64_t k = 0; k < nc; k++) {
      const uint32_t w = c.cigar[(size_t)(c0 + k)], op = w & 15u;
      if (op == 0 || op == 2 || op == 3 || op == 7 || op == 8) rlen += w >> 4;
    }
    const int32_t pos0 = c.pos[i] - 1;
    const int64_t b0 = pos0 > 0 ? pos0 : 0;
    const uint16_t f = c.flag[i];
    put32(out, (uint32_t)(32 + name.size() + 1 + 4 * nc + (L + 1) / 2 + L));
    put32(out, (uint32_t)ref);
    put32(out, (uint32_t)pos0);
    out.push_back((uint8_t)(name.size() + 1));
    out.push_back(c.mapq[i]);
    put16(out, (uint16_t)reg2bin(b0, b0 + (rlen > 1 ? rlen : 1)));
    put16(out, (uint16_t)nc);
    put16(out, (uint16_t)(f & 0x0FFFu));
    put32(out, (uint32_t)L);
    put32(out, (uint32_t)((f & PL_RF_MATE_SAME_REF) ? ref : -1));
    put32(out, (uint32_t)-1);
    put32(out, (uint32_t)c.tlen[i]);
    out.insert(out.end(), name.begin(), name.end());
    out.push_back(0);
    for (int64_t k = 0; k < nc; k++) put32(out, c.cigar[(size_t)(c0 + k)]);
    const bool exo = e < c.exo_read.size() && c.exo_read[e] == (int64_t)i;
    for (int64_t k = 0; k < L; k += 2) {
      uint8_t hi, lo = 0;
      if (exo) {
        const int64_t o = c.exo_off[e] + k;
        hi = (uint8_t)((c.nib[(size_t)(o >> 1)] >> (4 * (1 - (o & 1)))) & 15);
        if (k + 1 < L) lo = (uint8_t)((c.nib[(size_t)((o + 1) >> 1)] >> (4 * (1 - ((o + 1) & 1)))) & 15);
      } else {
        const int64_t q = s0 + k;
        hi = kNib[(c.bases2[(size_t)(q >> 2)] >> (2 * (q & 3))) & 3];
        if (k + 1 < L) lo = kNib[(c.bases2[(size_t)((q + 1) >> 2)] >> (2 * ((q + 1) & 3))) & 3];
      }
      out.push_back((uint8_t)((hi << 4) | lo));
    }
    if (exo) e++;
    if (f & PL_RF_NO_QUAL) out.insert(out.end(), (size_t)L, (uint8_t)0xFF);
    else out.insert(out.end(), (const uint8_t*)c.quals.data() + s0, (const uint8_t*)c.quals.data() + s0 + L);
  }
}
bool deflate_block(const uint8_t* src, size_t n, int level, std::vector<uint8_t>& out) {
  z_stream zs;
  std::memset(&zs, 0, sizeof zs);
  if (deflateInit2(&zs, level, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) return false;
  std::vector<uint8_t> cd(deflateBound(&zs, (uLong)n) + 16);
  zs.next_in = const_cast<Bytef*>(src);
  zs.avail_in = (uInt)n;
  zs.next_out = cd.data();
  zs.avail_out = (uInt)cd.size();
  const int rc = deflate(&zs, Z_FINISH);
  const size_t clen = zs.total_out;
  deflateEnd(&zs);
  if (rc != Z_STREAM_END) return false;
  const size_t bsize = 18 + clen + 8;
  const uint8_t head[18] = {31, 139, 8, 4, 0, 0, 0, 0, 0, 255, 6, 0, 66, 67, 2, 0,
                            (uint8_t)((bsize - 1) & 255), (uint8_t)((bsize - 1) >> 8)};
  out.assign(head, head + 18);
  out.insert(out.end(), cd.begin(), cd.begin() + (std::ptrdiff_t)clen);
  put32(out, (uint32_t)crc32(0L, src, (uInt)n));
  put32(out, (uint32_t)n);
  return true;
}
}  // namespace

extern "C" int64_t synth_write_bam(void* h, const char* path, const char* const* names, int32_t level) {
  SynthResult& S = *(SynthResult*)h;
  const int32_t nc = (int32_t)S.contigs.size();
  std::vector<uint8_t> stream;
  const std::string text = "@HD\tVN:1.6\tSO:coordinate\n";
  stream.insert(stream.end(), {'B', 'A', 'M', 1});
  put32(stream, (uint32_t)text.size());
  stream.insert(stream.end(), text.begin(), text.end());
  put32(stream, (uint32_t)nc);
  for (int32_t ci = 0; ci < nc; ci++) {
    const size_t ln = std::strlen(names[ci]);
    put32(stream, (uint32_t)(ln + 1));
    stream.insert(stream.end(), names[ci], names[ci] + ln);
    stream.push_back(0);
    put32(stream, (uint32_t)S.contigs[(size_t)ci].draft.size());
  }
  std::vector<std::vector<uint8_t>> rec((size_t)nc);
#pragma omp parallel for schedule(dynamic)
  for (int32_t ci = 0; ci < nc; ci++) contig_records(S.contigs[(size_t)ci], ci, rec[(size_t)ci]);
  size_t total = stream.size();
  for (auto& r : rec) total += r.size();
  stream.reserve(total);
  for (auto& r : rec) { stream.insert(stream.end(), r.begin(), r.end()); std::vector<uint8_t>().swap(r); }
  const size_t B = 0xFF00, nb = (stream.size() + B - 1) / B;
  std::vector<std::vector<uint8_t>> blk(nb + 1);
  int ok = 1;
#pragma omp parallel for schedule(static) reduction(&& : ok)
  for (int64_t k = 0; k < (int64_t)nb; k++) {
    const size_t o = (size_t)k * B;
    ok = deflate_block(stream.data() + o, std::min(B, stream.size() - o), level, blk[(size_t)k]) && ok;
  }
  if (!ok || !deflate_block(nullptr, 0, level, blk[nb])) return -1;
  FILE* f = std::fopen(path, "wb");
  if (!f) return -1;
  for (auto& b : blk) if (std::fwrite(b.data(), 1, b.size(), f) != b.size()) { std::fclose(f); return -1; }
  std::fclose(f);
  return (int64_t)stream.size();
}

