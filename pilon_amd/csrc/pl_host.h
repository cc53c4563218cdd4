// pl_host.h — host-side pieces of the path that act on the sparse change
// list: final dispositions, identifyAndFixIssues (snps/indels), fixFixList,
// fixIssues and writeChanges.  Header-only C++17.
//
// Reference: broadinstitute/pilon v1.24, src/main/scala/org/broadinstitute/pilon/
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include "pl_common.h"

namespace plh {

// A position's disposition after Pass 1: the window kernel stores the
// candidate assuming depth >= minDepth and not deleted; these are the two
// remaining tests of GenomeRegion.scala:255.  No Pass 1 runs when the region
// has no reads (GenomeRegion.scala:229-231).
inline uint16_t final_flags(uint16_t flags, int64_t depth, int32_t min_depth, int64_t read_count) {
  const int cand = (flags & PL_CF_DISP_MASK) >> PL_CF_DISP_SHIFT;
  int disp = cand;
  if (read_count == 0 || (flags & PL_CF_DELETED) || depth < min_depth) disp = PL_D_NONE;
  return (uint16_t)((flags & ~PL_CF_DISP_MASK) | (disp << PL_CF_DISP_SHIFT));
}

struct Change {
  int32_t loc;        // 1-based locus
  int kind;           // PL_D_SNP / INS / DEL / AMB
  uint8_t ref;        // refBase (upper case)
  uint8_t base, alt;  // BaseCall.base / altBase codes (0..3, 4 = N)
  std::string indel;  // BaseCall.insertion / deletion
};

struct Fix { int32_t loc; std::string was, patch; };
// one logged reassembly attempt (GenomeRegion.logFix): its region, the
// reassemblyFixes outcome and the log line
struct Outcome { int32_t start, stop; std::string outcome, line; };

// GenomeRegion.excludeMotifs (GenomeRegion.scala:219,416-439): when every BAM
// of a long-read run is a non-frags one (longReadOnly, :147), identifyAndFixIssues
// neither fixes nor counts a change (:326) at a locus that starts a run of 4
// equal upper-cased bases within the region, or, with a nanopore BAM, that
// sits in the middle of CC_GG.  contig holds the whole contig (0-based).
inline void exclude_motifs(std::vector<Change>& ch, const uint8_t* contig, int32_t start, int32_t stop, bool nano) {
  auto at = [&](int64_t loc) { const uint8_t c = contig[loc - 1]; return (c >= 'a' && c <= 'z') ? (uint8_t)(c - 32) : c; };
  std::vector<Change> keep;
  keep.reserve(ch.size());
  for (Change& x : ch) {
    const int64_t loc = x.loc;
    int64_t run = 1;                                         // homoRun(loc), bounded by the region's stop
    while (loc + run <= stop && run < 4 && at(loc + run) == at(loc)) run++;
    const bool motif = nano && loc - 2 >= start && loc + 2 <= stop && at(loc - 2) == 'C' && at(loc - 1) == 'C' &&
                       at(loc + 1) == 'G' && at(loc + 2) == 'G';
    if (run < 4 && !motif) keep.push_back(std::move(x));
  }
  ch.swap(keep);
}

inline char iupac(int b1, int b2) {               // Bases.toIUPAC (Bases.scala:85-88)
  static const char T[16] = {0, 'A', 'C', 'M', 'G', 'R', 'S', 'V', 'T', 'W', 'Y', 'H', 'K', 'D', 'B', 'N'};
  return T[(1 << b1) | (1 << b2)];
}

// GenomeRegion.fixFixList (GenomeRegion.scala:557-593): stable sort by locus,
// of overlapping neighbours keep the one with the larger was+patch, ties to
// the earlier.
inline std::vector<Fix> fix_fix_list(std::vector<Fix> in) {
  std::stable_sort(in.begin(), in.end(), [](const Fix& a, const Fix& b) { return a.loc < b.loc; });
  std::vector<Fix> out;
  if (in.empty()) return out;
  Fix cur = in[0];
  for (size_t i = 1; i < in.size(); i++) {
    const Fix& nx = in[i];
    const int64_t e1 = (int64_t)cur.loc + std::max<int64_t>((int64_t)cur.was.size() - 1, 0);
    const int64_t e2 = (int64_t)nx.loc + std::max<int64_t>((int64_t)nx.was.size() - 1, 0);
    if ((int64_t)nx.loc <= e1 && e2 >= (int64_t)cur.loc) {
      if (cur.was.size() + cur.patch.size() < nx.was.size() + nx.patch.size()) cur = nx;
    } else {
      out.push_back(cur);
      cur = nx;
    }
  }
  out.push_back(cur);
  return out;
}

// GenomeRegion.fixIssues (GenomeRegion.scala:595-621).  Pilon applies the
// kept fixes right to left; they are sorted and no two overlap (each starts
// after the last base its predecessor replaces), so one left-to-right splice
// gives the same bases in time linear in the region.
inline void apply_fixes(std::string& bases, int32_t region_start, const std::vector<Fix>& list) {
  const std::vector<Fix> kept = fix_fix_list(list);
  if (kept.empty()) return;
  std::string out;
  out.reserve(bases.size() + 64);
  size_t at = 0;
  for (const Fix& f : kept) {
    const size_t s = (size_t)(f.loc - region_start);
    out.append(bases, at, s - at);
    out += f.patch;
    at = s + f.was.size();
  }
  out.append(bases, at, std::string::npos);
  bases.swap(out);
}

// GenomeFile.processRegions' newName (GenomeFile.scala:137-140): name + sep +
// "pilon", sep "_" unless the name holds a '|' ("" when it ends with one).
inline std::string pilon_name(const std::string& name) {
  const char* sep = name.find('|') == std::string::npos ? "_" : (!name.empty() && name.back() == '|') ? "" : "|";
  return name + sep + "pilon";
}

inline std::string region_string(const std::string& name, int64_t s, int64_t e) {  // Region.regionString
  char buf[64];
  if (e + 1 - s < 2) snprintf(buf, sizeof buf, ":%lld", (long long)s);
  else snprintf(buf, sizeof buf, ":%lld-%lld", (long long)s, (long long)e);
  return name + buf;
}

struct FixResult { std::string changes, bases; int snps = 0, amb = 0, ins = 0, ins_bases = 0, dels = 0, del_bases = 0; };

// identifyAndFixIssues' snpFixList and smallFixList (GenomeRegion.scala:
// 307-355), each in List order (the last one prepended first), with the
// counts it logs.  changes must be sorted by locus.
inline void fix_lists(const pl_config& cf, const std::vector<Change>& changes, std::vector<Fix>& snp,
                      std::vector<Fix>& small, FixResult& R) {
  for (const Change& c : changes) {
    const std::string r(1, (char)c.ref);
    const std::string b(1, c.base == 4 ? 'N' : "ACGT"[c.base]);
    switch (c.kind) {
      case PL_D_SNP:
        if (cf.fix_snps) snp.push_back({c.loc, r, b});
        R.snps++;
        break;
      case PL_D_AMB:
        if (cf.fix_snps && !cf.longread) {
          if (cf.iupac) small.push_back({c.loc, r, std::string(1, iupac(c.base, c.alt))});
          else snp.push_back({c.loc, r, b});
          R.amb++;
        }
        break;
      case PL_D_INS:
        if (cf.fix_indels) small.push_back({c.loc, "", c.indel});
        R.ins++;
        R.ins_bases += (int)c.indel.size();
        break;
      case PL_D_DEL:
        if (cf.fix_indels) small.push_back({c.loc, c.indel, ""});
        R.dels++;
        R.del_bases += (int)c.indel.size();
        break;
    }
  }
  std::reverse(snp.begin(), snp.end());
  std::reverse(small.begin(), small.end());
}

// The region's bases as local reassembly sees them: after
// fixIssues(snpFixList) (GenomeRegion.scala:380).
inline std::string snp_fixed_bases(const pl_config& cf, int32_t start, const std::string& region_bases,
                                   const std::vector<Change>& changes) {
  std::vector<Fix> snp, small;
  FixResult R;
  fix_lists(cf, changes, snp, small, R);
  std::string b = region_bases;
  apply_fixes(b, start, snp);
  return b;
}

// identifyAndFixIssues for one region (GenomeRegion.scala:307-413): the
// snps/indels fixes, with big (bigFixList, in List order: the reassembly
// fixes, the last one found first) applied after them, and writeChanges
// (:645-657).  region_bases is contigBases[start-1 .. stop-1].
inline FixResult fix_region(const pl_config& cf, const std::string& name, int32_t start,
                            const std::string& region_bases, const std::vector<Change>& changes,
                            const std::string& new_name, int64_t offset, const std::vector<Fix>& big = {}) {
  FixResult R;
  std::vector<Fix> snp, small;
  fix_lists(cf, changes, snp, small, R);
  R.bases = region_bases;
  apply_fixes(R.bases, start, snp);
  std::vector<Fix> later = small;                  // fixIssues(smallFixList ++ bigFixList)
  later.insert(later.end(), big.begin(), big.end());
  apply_fixes(R.bases, start, later);
  std::vector<Fix> all = snp;
  all.insert(all.end(), later.begin(), later.end());
  int64_t delta = 0;
  for (const Fix& f : fix_fix_list(all)) {
    const int64_t newLoc = f.loc + delta;
    R.changes += region_string(name, f.loc, (int64_t)f.loc + (int64_t)f.was.size() - 1);
    R.changes += ' ';
    R.changes += region_string(new_name, newLoc + offset, newLoc + offset + (int64_t)f.patch.size() - 1);
    R.changes += ' ';
    R.changes += f.was.empty() ? std::string(".") : f.was;
    R.changes += ' ';
    R.changes += f.patch.empty() ? std::string(".") : f.patch;
    R.changes += '\n';
    delta += (int64_t)f.patch.size() - (int64_t)f.was.size();
  }
  return R;
}

}  // namespace plh
