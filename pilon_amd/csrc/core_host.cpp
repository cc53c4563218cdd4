// core_host.cpp — host build of the kernel core (pl_core.h), for CPU tests only.
//
// It runs the same per-read walk, window flush, indel pass and Pass-1
// finalize the gfx950 kernels run, one window and one read at a time, so the
// window decomposition (scan margins, outliers, deferral of indel columns,
// deletion pass) can be compared with the oracle without a GPU.  It is not
// part of libpilonhip.so and the product never loads it.
#include <algorithm>
#include <array>
#include <map>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pl_core.h"
#include "pl_host.h"

namespace {

struct HostSink {
  int W;
  uint64_t *bs, *mqq, *ins_, *del_;
  uint32_t *bad_, *clp_;
  int32_t *pc, *is;
  std::vector<IndelRec>* recs;
  int64_t bcount = 0, aligned_n = 0, edge_read = -1;
  void frag(int p, int d, int32_t size) { pc[p] = j_add(pc[p], d); is[p] = j_add(is[p], size); }
  void aligned(int n) { aligned_n += n; }
  void base(int p, int b, int32_t q, int32_t mq1) { bs[b * W + p] += pack_base(q, mq1); mqq[p] += pack_mq_q(mq1, q); }
  void count_base() { bcount++; }
  void bad(int p, int n) { bad_[p] += (uint32_t)n; }
  void clip(int p) { clp_[p] += 1; }
  void ins(int p, int32_t q, int32_t mq1, const IndelRec& e) {
    ins_[p] += (1ull << 32) + (uint64_t)mq1;
    mqq[p] += (uint64_t)(int64_t)q << 32;
    recs->push_back(e);
  }
  void del(int p, int32_t q, int32_t mq1, const IndelRec& e) {
    del_[p] += (1ull << 32) + (uint64_t)mq1;
    mqq[p] += pack_mq_q(mq1, q);
    recs->push_back(e);
  }
  void edge(int64_t r) { if (edge_read < 0 || r < edge_read) edge_read = r; }
};

struct Handle {
  std::vector<std::string> changes, bases;
};

int64_t lower_pos(const int32_t* pos, int64_t n, int64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) { int64_t m = (lo + hi) / 2; if (pos[m] < x) lo = m + 1; else hi = m; }
  return lo;
}

}  // namespace

extern "C" {

// Runs n_jobs regions of one contig with window size W.  pile (PL_NFIELDS x
// G) and calls (G) are written in job order; stats/errs have n_jobs entries.
void* ch_run(const pl_config* pc, const uint8_t* ref, int64_t ref_len, const char* name,
             const pl_read_soa* s, int32_t n_jobs, const int32_t* starts, const int32_t* stops,
             int32_t W, int32_t* pile_out, pl_call* calls_out, pl_region_stats* stats_out,
             int32_t* errs, const uint8_t* lrt, int32_t set_kinds) {
  // lrt: longReadType per read (null: none); set_kinds bit 0: a frags BAM
  // is among the set's, bit 1: a nanopore BAM is
  KConf cf;
  cf.flank = pc->flank; cf.min_mq = pc->min_mq; cf.min_qual = pc->min_qual;
  cf.default_qual = pc->default_qual; cf.min_min_depth = pc->min_min_depth;
  cf.old_indel = pc->old_indel; cf.fixamb = pc->iupac || pc->fix_amb;
  cf.nonpf = pc->nonpf; cf.duplicates = pc->duplicates; cf.min_depth = pc->min_depth;
  const int64_t n = s->n_reads;
  std::vector<RMeta> meta((size_t)n + 1);
  RSDev R;
  std::memset(&R, 0, sizeof R);
  R.n = n; R.pos = s->pos; R.tlen = s->tlen; R.flag = s->flag; R.mapq = s->mapq;
  R.cigar_off = s->cigar_off; R.cigar = s->cigar; R.seq_off = s->seq_off;
  R.bases2 = s->bases2; R.quals = s->quals; R.n_exotic = s->n_exotic;
  R.exo_read = s->exo_read; R.exo_off = s->exo_off; R.nib = s->nib; R.lrt = lrt;
  R.meta = meta.data(); R.ref = ref; R.ref_len = ref_len;
  // ---- read-set prep (k_readset_prep)
  int32_t max_back = 0, max_lead = 0, max_back_fp = 0;
  std::vector<int64_t> outliers;
  for (int64_t r = 0; r < n; r++) meta[r] = prep_read(R, r, cf).m;
  // segment table (k_prep counts, k_segfill fills)
  std::vector<Seg> segs;
  for (int64_t r = 0; r < n; r++) {
    RMeta& m = meta[r];
    if (!((m.info & RI_KEEP) && !(m.info & (RI_SIMPLE | RI_EXOTIC | RI_BAD | RI_LR_MASK)))) continue;
    const int c = build_segs(R, r, m.info, cf, nullptr);
    const size_t o = segs.size();
    segs.resize(o + (size_t)c);
    if (c) build_segs(R, r, m.info, cf, segs.data() + o);
    m.op0 = (uint32_t)o;
    m.info = (m.info & ~RI_SEG_MASK) | ((uint32_t)(c < RI_SEG_MAX ? c : RI_SEG_MAX) << RI_SEG_SHIFT);
  }
  segs.emplace_back();
  R.segs = segs.data();
  for (int64_t r = 0; r < n; r++) {
    const RMeta& m = meta[r];
    if (!(m.info & RI_KEEP)) continue;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    if (m.fp_lo <= m.fp_hi) { lo = m.fp_lo; hi = m.fp_hi; }
    if (m.fr_lo < m.fr_hi) { lo = std::min<int64_t>(lo, m.fr_lo); hi = std::max<int64_t>(hi, (int64_t)m.fr_hi - 1); }
    if (lo > hi) continue;
    int64_t lead = R.pos[r] - lo, back = hi - R.pos[r];
    if (lead > PL_MAX_LEAD || back > PL_MAX_BACK || R.seq_off[r + 1] - R.seq_off[r] >= PL_MAX_SEQ) { outliers.push_back(r); meta[r].info |= RI_OUTLIER; continue; }
    max_back = std::max<int32_t>(max_back, (int32_t)std::max<int64_t>(back, 0));
    max_lead = std::max<int32_t>(max_lead, (int32_t)std::max<int64_t>(lead, 0));
    if (m.fp_lo <= m.fp_hi && m.fp_hi > m.pos) max_back_fp = std::max<int32_t>(max_back_fp, m.fp_hi - m.pos);
  }
  // ---- jobs and windows
  int64_t G = 0;
  std::vector<JobDev> jobs((size_t)n_jobs);
  std::vector<JobStat> st((size_t)n_jobs);
  for (int32_t j = 0; j < n_jobs; j++) {
    JobDev& J = jobs[j];
    std::memset(&J, 0, sizeof J);
    J.rs = 0; J.start = starts[j]; J.stop = stops[j];
    J.qs = (int32_t)std::max<int64_t>((int64_t)J.start - 10000, 0);
    J.qe = (int32_t)std::min<int64_t>((int64_t)J.stop + 10000, ref_len);
    J.size = J.stop - J.start + 1;
    J.out = G;
    G += J.size;
    std::memset(&st[j], 0, sizeof(JobStat));
    st[j].err_read_edge = INT64_MAX;
    st[j].err_read_inval = INT64_MAX;
  }
  int32_t* pile = pile_out;
  pl_call* calls = calls_out;
  std::vector<IndelRec> recs;
  std::vector<uint32_t> defer;
  std::map<uint32_t, std::array<int64_t, 4>> wide;   // exact qualSums of wide columns
  auto patch = [&](Col& c, uint32_t g) {
    auto it = wide.find(g);
    if (it != wide.end()) for (int b = 0; b < 4; b++) c.qs[b] = it->second[b];
  };
  std::vector<Cand> cands;
  std::vector<uint64_t> bs(4 * (size_t)W), mqq(W), insA(W), delA(W);
  std::vector<uint32_t> badA(W), clpA(W);
  std::vector<int32_t> pcA(W), isA(W);
  for (int32_t j = 0; j < n_jobs; j++) {
    const JobDev& J = jobs[j];
    JobStat& S = st[j];
    // count (k_jobprep): kept reads overlapping the query
    for (int64_t r = 0; r < n; r++) {
      uint32_t info = meta[r].info;
      if (!(info & RI_KEEP)) continue;
      int32_t spanEnd = meta[r].aend != 0 ? meta[r].aend : R.pos[r];
      if (!(R.pos[r] <= J.qe && spanEnd >= J.qs)) continue;
      S.read_count++;
      S.unknown_ops += (info & RI_UNK_MASK) >> RI_UNK_SHIFT;
      if ((info & RI_BAD) && r < S.err_read_inval) S.err_read_inval = r;
      if ((info & RI_EDGE_CLIP) && r < S.err_read_edge) S.err_read_edge = r;
    }
    for (int32_t p0 = 0; p0 < J.size; p0 += W) {
      const int32_t p1 = std::min<int32_t>(J.size, p0 + W);
      const int32_t L0 = J.start + p0, L1 = J.start + p1;
      std::fill(bs.begin(), bs.end(), 0); std::fill(mqq.begin(), mqq.end(), 0);
      std::fill(insA.begin(), insA.end(), 0); std::fill(delA.begin(), delA.end(), 0);
      std::fill(badA.begin(), badA.end(), 0); std::fill(clpA.begin(), clpA.end(), 0);
      std::fill(pcA.begin(), pcA.end(), 0); std::fill(isA.begin(), isA.end(), 0);
      HostSink sk;
      sk.W = W; sk.bs = bs.data(); sk.mqq = mqq.data(); sk.ins_ = insA.data(); sk.del_ = delA.data();
      sk.bad_ = badA.data(); sk.clp_ = clpA.data(); sk.pc = pcA.data(); sk.is = isA.data(); sk.recs = &recs;
      const int64_t lo = lower_pos(R.pos, n, (int64_t)L0 - max_back);
      const int64_t hi = lower_pos(R.pos, n, (int64_t)L1 + max_lead);
      // fragment-only reads from their tails, as the window kernel does
      const int64_t plo = fragment_only_end(R.pos, lo, hi, J, L0, max_back, max_back_fp);
      int32_t c0 = 0, s0 = 0;
      for (int64_t r = lo; r < plo; r++) {
        const FragTail t = frag_tail(meta[r]);
        if (frag_tail_event(t, L0, L1, sk)) { c0++; s0 = j_add(s0, t.size); }
      }
      if (c0) sk.frag(0, c0, s0);
      // same staging as the window kernel: runs for the flattened loop,
      // one-wave walks for the rest
      std::vector<Run> runs;
      std::vector<int64_t> cplx;
      for (int64_t r = plo; r < hi; r++) {
        const RMeta m = meta[r];
        if ((m.info & RI_OUTLIER) || !read_in_query(m, J)) continue;
        frag_event(m, J, L0, L1, sk);
        if (m.fp_hi < L0 || m.fp_lo >= L1) continue;
        Run rr[1];
        const int ns = (int)((m.info & RI_SEG_MASK) >> RI_SEG_SHIFT);
        if (m.info & RI_SIMPLE) {
          if (simple_run(m, R.seq_off[r], cf, L0, L1, rr, &sk.aligned_n, &sk.bcount)) runs.push_back(rr[0]);
        } else if (!(m.info & RI_EXOTIC) && ns < RI_SEG_MAX) {
          // each segment on its own, as the kernel's lanes apply them
          const int32_t L = (int32_t)(R.seq_off[r + 1] - R.seq_off[r]);
          for (int i = 0; i < ns; i++)
            if (seg_apply(J, cf, r, m.info, R.seq_off[r], L, L0, L1, R.segs[m.op0 + i], rr[0], &sk.aligned_n, &sk.bcount, sk))
              runs.push_back(rr[0]);
        } else {
          cplx.push_back(r);
        }
      }
      for (const Run& ru : runs)
        for (int32_t i = 0; i < ru.n; i++) {
          const int p = ru.p0 + i;
          if (ru.fl & EF_VALID) {
            const int64_t q = ru.q + i;
            const int32_t qual = (ru.fl & EF_NOQUAL) ? cf.default_qual : (int32_t)R.quals[q];
            const int b = (R.bases2[q >> 2] >> (2 * (q & 3))) & 3;
            if (qual >= cf.min_qual) sk.base(p, b, qual, (int32_t)((ru.fl >> 16) & 0x1ffu));
          } else {
            sk.bad(p, 1);
          }
        }
      for (int64_t r : cplx) walk_read(R, J, cf, r, L0, L1, 0, 1, sk, false);
      for (int64_t r : outliers) walk_read(R, J, cf, r, L0, L1, 0, 1, sk, true);
      if (sk.edge_read >= 0 && sk.edge_read < S.err_read_edge) S.err_read_edge = sk.edge_read;
      S.base_count += sk.bcount;
      S.aligned += sk.aligned_n;
      int32_t apc = 0, ais = 0;
      for (int32_t p = 0; p < p1 - p0; p++) {
        apc = j_add(apc, pcA[p]);
        ais = j_add(ais, isA[p]);
        uint64_t b4[4] = {bs[0 * W + p], bs[1 * W + p], bs[2 * W + p], bs[3 * W + p]};
        const int64_t g = J.out + p0 + p;
        PosOut o;
        finalize_position(b4, mqq[p], insA[p], delA[p], badA[p], clpA[p], apc, ais, ref[L0 + p - 1], cf, o);
        if (o.overflow) S.overflow = 1;
        if (o.wide) {
          o.call.flags |= PL_CF_WIDE;
          wide[(uint32_t)g] = {o.qs[0], o.qs[1], o.qs[2], o.qs[3]};
        }
        for (int f = 0; f < PL_NFIELDS; f++) pile[(int64_t)f * G + g] = o.f[f];
        calls[g] = o.call;
        if (o.defer) defer.push_back((uint32_t)g);
        else {
          const int d = call_disp(o.call);
          if (d == PL_D_CONFIRMED) {
            int64_t dep = 0;
            for (int b = 0; b < 4; b++) dep += (uint32_t)o.f[PL_F_COUNT_A + b];
            dep += o.f[PL_F_DELETIONS];
            S.hist[dep > 255 ? 255 : dep]++;
          } else if (d == PL_D_SNP || d == PL_D_AMB) {
            int64_t dep = 0;
            for (int b = 0; b < 4; b++) dep += (uint32_t)o.f[PL_F_COUNT_A + b];
            dep += o.f[PL_F_DELETIONS];
            cands.push_back({(uint32_t)g, (int32_t)dep, d, 0, -1});
          }
        }
      }
    }
  }
  // ---- sort records by key (stable), indel pass on deferred columns
  std::vector<uint32_t> rkey(recs.size()), rval(recs.size());
  for (size_t i = 0; i < recs.size(); i++) rval[i] = (uint32_t)i;
  std::stable_sort(rval.begin(), rval.end(), [&](uint32_t a, uint32_t b) { return recs[a].key < recs[b].key; });
  for (size_t i = 0; i < recs.size(); i++) rkey[i] = recs[rval[i]].key;
  auto job_of = [&](int64_t g) {
    int32_t j = 0;
    while (j + 1 < n_jobs && jobs[j + 1].out <= g) j++;
    return j;
  };
  for (uint32_t g : defer) {
    const int32_t j = job_of(g);
    const JobDev& J = jobs[j];
    Col c = load_col(pile, G, g);
    patch(c, g);
    int32_t ilen;
    int64_t wrec;
    pl_call cl = full_call(&R, J, c, ref[J.start + (g - J.out) - 1], recs.data(), rkey.data(), rval.data(),
                           (int64_t)recs.size(), g, cf, &ilen, &wrec);
    cl.flags |= calls[g].flags & PL_CF_WIDE;
    calls[g] = cl;
    const int d = call_disp(cl);
    const int64_t dep = c.depth();
    if (d == PL_D_CONFIRMED) st[j].hist[dep > 255 ? 255 : dep]++;
    else if (d != PL_D_NONE) cands.push_back({g, (int32_t)dep, d, ilen, wrec});
  }
  std::stable_sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.gpos < b.gpos; });
  std::vector<uint32_t> mutated((size_t)G + 1);
  uint32_t mut_n = 0;
  int32_t ovf = 0;
  Handle* H = new Handle();
  for (int32_t j = 0; j < n_jobs; j++) {
    const JobDev& J = jobs[j];
    int64_t a = 0, b = 0;
    while (a < (int64_t)cands.size() && cands[a].gpos < (uint64_t)J.out) a++;
    b = a;
    while (b < (int64_t)cands.size() && cands[b].gpos < (uint64_t)(J.out + J.size)) b++;
    uint32_t m0 = mut_n;
    finalize_job(cf, J, st[j], pile, G, calls, cands.data(), a, b, mutated.data(), &mut_n, (uint32_t)mutated.size(), &ovf);
    for (uint32_t k = m0; k < mut_n; k++) {
      const uint32_t g = mutated[k];
      Col c = load_col(pile, G, g);
      patch(c, g);
      int32_t ilen;
      int64_t wrec;
      pl_call cl = full_call(&R, J, c, ref[J.start + (g - J.out) - 1], recs.data(), rkey.data(), rval.data(),
                             (int64_t)recs.size(), g, cf, &ilen, &wrec);
      cl.flags |= calls[g].flags & PL_CF_WIDE;
      calls[g].base = cl.base; calls[g].alt = cl.alt; calls[g].q = cl.q;
      calls[g].flags = (uint16_t)((cl.flags & ~PL_CF_DISP_MASK) | (calls[g].flags & PL_CF_DISP_MASK) | PL_CF_DELETED);
    }
    // final dispositions and the change list
    std::vector<plh::Change> ch;
    for (int32_t p = 0; p < J.size; p++) {
      const int64_t g = J.out + p;
      int64_t dep = 0;
      for (int b2 = 0; b2 < 4; b2++) dep += (uint32_t)pile[(int64_t)(PL_F_COUNT_A + b2) * G + g];
      dep += pile[(int64_t)PL_F_DELETIONS * G + g];
      calls[g].flags = plh::final_flags(calls[g].flags, dep, st[j].min_depth, st[j].read_count);
      if (st[j].min_depth > PL_CONF_HIST_MAX && call_disp(calls[g]) == PL_D_CONFIRMED) st[j].n_confirmed++;
    }
    for (int64_t k = a; k < b; k++) {
      const Cand& c = cands[k];
      if (!c.kind) continue;
      plh::Change x;
      const int32_t p = (int32_t)(c.gpos - J.out);
      x.loc = J.start + p;
      x.kind = c.kind;
      x.ref = to_upper(ref[x.loc - 1]);
      x.base = calls[c.gpos].base;
      x.alt = calls[c.gpos].alt;
      if (c.kind == PL_D_INS) {
        const IndelRec& e = recs[c.rec];
        for (int32_t q = 0; q < e.len; q++) x.indel.push_back((char)rec_char(R, e, q));
      } else if (c.kind == PL_D_DEL) {
        x.indel.assign((const char*)ref + x.loc - 1, (size_t)c.len);
      }
      ch.push_back(x);
    }
    if (pc->longread && !(set_kinds & 1)) plh::exclude_motifs(ch, ref, J.start, J.stop, (set_kinds & 2) != 0);
    std::string nm(name);
    std::string rb((const char*)ref + J.start - 1, (size_t)J.size);
    plh::FixResult fr = plh::fix_region(*pc, nm, J.start, rb, ch, nm + "_pilon", 0);
    H->changes.push_back(fr.changes);
    H->bases.push_back(fr.bases);
    pl_region_stats& o = stats_out[j];
    std::memset(&o, 0, sizeof o);
    o.read_count = st[j].read_count;
    o.base_count = st[j].base_count;
    o.coverage = round_div_l(st[j].base_count, J.size);
    o.min_depth = st[j].min_depth;
    o.n_confirmed = st[j].n_confirmed;
    o.snps = fr.snps; o.amb = fr.amb; o.ins = fr.ins; o.ins_bases = fr.ins_bases;
    o.dels = fr.dels; o.del_bases = fr.del_bases;
    o.n_unknown_ops = (int32_t)st[j].unknown_ops;
    o.aligned_bases = st[j].aligned;
    errs[j] = st[j].err_read_inval < st[j].err_read_edge ? PL_E_INVAL
              : st[j].err_read_edge != INT64_MAX ? PL_E_EDGE : st[j].overflow ? PL_E_OVERFLOW : PL_OK;
  }
  (void)ovf;
  return H;
}

int64_t ch_text(void* h, int32_t j, int32_t which, const char** p) {
  Handle* H = (Handle*)h;
  const std::string& s = which == 0 ? H->changes[j] : H->bases[j];
  *p = s.data();
  return (int64_t)s.size();
}

void ch_free(void* h) { delete (Handle*)h; }

}  // extern "C"
