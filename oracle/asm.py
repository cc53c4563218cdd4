"""Oracle (test infrastructure only): local reassembly across a gap or a
break, restated from the reference in plain Python.  Only tests/ import it.

The K-mer graph comes from the oracle's own Assembler.addReads + buildGraph
(`orc_kmer_graph`, oracle.kmer_graph).  On it this module restates
Assembler's path search (Assembler.scala:163-253: kmerPathString,
noteKmerLoop, kmerPathsForward, pathsForward, tryForward, tryReverse,
multiBridge) and GapFiller's assembly across the break (GapFiller.scala:
39-262: assembleAcrossBreak, consensusFromLeft/Right, assembleIntoBreak,
breakJoins, joinBreak, properOverlap, trimPatch, partialMatchesReference).

Pilon keeps the joins in a Scala Set and lists them in its iteration order;
up to four elements that is insertion order, which is what this restatement
uses (beyond four distinct joins the order is that of a hash set: parity
unpinned there).  A `require` that Pilon would fail (a substring outside the
region) raises Edge.
"""

MIN_EXTEND = 20      # Assembler.minExtend and GapFiller.minExtend (Assembler.scala:27, GapFiller.scala:29)
MAX_BRANCHES = 5     # Assembler.maxBranches (Assembler.scala:28)
NO_SOLUTION = (0, "", "")


class Edge(Exception):
    """An input on which Pilon itself throws."""


def complement(c):                                   # Bases.complement (Bases.scala:47-53)
    return {"A": "T", "C": "G", "G": "C", "T": "A"}.get(c, c)


def reverse_complement(s):                           # Bases.reverseComplement (Bases.scala:58-60)
    return "".join(complement(c) for c in reversed(s))


class Assembler:
    """The graph (kGraph / altGraph, Assembler.scala:37-39,131-153) and the
    path search, with the loop Pilon notes across all searches."""

    def __init__(self, graph, K, tr_safe=True):
        self.K = K
        self.tr_safe = tr_safe
        self.kgraph, self.altgraph = {}, {}
        for kmer, base, alt in graph:
            prefix = kmer[1:]
            self.kgraph[kmer] = prefix + base
            if alt is not None:
                self.altgraph[kmer] = prefix + alt
        self.loop_length = 0
        self.loop_sequence = ""

    @staticmethod
    def kmer_path_string(kmers):                     # Assembler.scala:163-170 (kmers: newest first)
        path = list(reversed(kmers))
        return path[0] + "".join(k[-1] for k in path[1:])

    def note_kmer_loop(self, loop_index, kmers):     # Assembler.scala:172-180
        length = loop_index + 1
        if self.loop_length == 0 or length < self.loop_length:
            self.loop_length = length
            self.loop_sequence = self.kmer_path_string(kmers[:loop_index + 1])[:length]

    def kmer_paths_forward(self, kmers, branches=0):  # Assembler.scala:182-233 (kmers: newest first)
        kmers = list(kmers)
        while True:
            kmer = kmers[0]
            if kmer not in self.kgraph:
                return [kmers]
            if kmer in self.altgraph:
                next1, next2 = self.kgraph[kmer], self.altgraph[kmer]
                seen1, seen2 = next1 in kmers[1:], next2 in kmers[1:]
                if seen1 or seen2:
                    loop = max(kmers.index(next1) if next1 in kmers else -1,
                               kmers.index(next2) if next2 in kmers else -1)
                    self.note_kmer_loop(loop, kmers)
                    if self.tr_safe:
                        return [kmers]
                if seen1 and seen2:
                    return [kmers]
                elif seen1 and not seen2:
                    kmers.insert(0, next2)
                elif seen2 and not seen1:
                    kmers.insert(0, next1)
                else:
                    if branches < MAX_BRANCHES:
                        return self.kmer_paths_forward([next1] + kmers, branches + 1) + \
                            self.kmer_paths_forward([next2] + kmers, branches + 1)
                    return [kmers]
            else:
                nxt = self.kgraph[kmer]
                loop = kmers.index(nxt) if nxt in kmers else -1
                if loop >= 0:
                    self.note_kmer_loop(loop, kmers)
                    if self.tr_safe:
                        return [kmers]
                    if kmers.count(nxt) > 1:
                        return [kmers]
                kmers.insert(0, nxt)

    def paths_forward(self, starting_kmer):          # Assembler.scala:235-244
        assert len(starting_kmer) == self.K
        paths = [self.kmer_path_string(p) for p in self.kmer_paths_forward([starting_kmer])]
        return sorted(paths, key=lambda p: -len(p))  # sortWith(a.length > b.length), stable

    def try_forward(self, anchor):                   # Assembler.scala:251-264
        if len(anchor) < self.K:
            return [anchor]
        starting = anchor[:self.K]
        paths = self.paths_forward(starting)
        if paths and len(paths[0]) > len(anchor) + MIN_EXTEND:
            return paths
        return [starting + p for p in self.try_forward(anchor[self.K:])]

    def try_reverse(self, anchor):                   # Assembler.scala:266-274
        return [reverse_complement(p) for p in self.try_forward(reverse_complement(anchor))]

    def multi_bridge(self, left, right):             # Assembler.scala:276-280
        forward = self.try_forward(left)
        reverse = self.try_reverse(right)
        return forward, reverse, self.loop_sequence


class Region:
    """What GapFiller reads of its GenomeRegion: the current (SNP-fixed)
    bases of the region and the contig as read (GenomeRegion.scala:35-37,
    766-781)."""

    def __init__(self, start, stop, bases, contig):
        assert len(bases) == stop + 1 - start
        self.start, self.stop, self.bases, self.contig = start, stop, bases, contig

    def in_region(self, locus):
        return self.start <= locus <= self.stop

    def base_at(self, locus):
        if not self.in_region(locus):
            raise Edge("baseAt outside the region")
        return self.bases[locus - self.start].upper()

    def original_base_at(self, locus):
        if not self.in_region(locus):
            raise Edge("originalBaseAt outside the region")
        return self.contig[locus - 1].upper()

    def sub_string(self, locus, n):
        if not (self.in_region(locus) and self.in_region(locus + n - 1)):
            raise Edge("subString outside the region")
        return self.bases[locus - self.start:locus - self.start + n].upper()

    def ref_sub_string(self, locus, n):
        if not (self.in_region(locus) and self.in_region(locus + n - 1)):
            raise Edge("refSubString outside the region")
        return self.contig[locus - 1:locus + n - 1].upper()


class GapFiller:
    def __init__(self, region, graph, K=47, break_radius=None, tr_safe=True, multi_closure=False,
                 fix_breaks=False, gap_margin=100000, min_gap=10):
        self.region = region
        self.K = K
        self.k = 2 * K + 1                           # GapFiller.k (GapFiller.scala:30)
        self.graph = graph
        self.break_radius = break_radius if break_radius is not None else 3 * K
        self.tr_safe, self.multi_closure, self.fix_breaks = tr_safe, multi_closure, fix_breaks
        self.gap_margin, self.min_gap = gap_margin, min_gap
        self.tandem_repeat = ""

    def assemble_across_break(self, bstart, bstop, is_gap):   # GapFiller.scala:49-98
        start, left_paths, right_paths, stop, loop = self.assemble_into_break(bstart, bstop)
        self.tandem_repeat = loop
        solutions = self.break_joins(start, left_paths, right_paths, stop)
        if len(solutions) == 1 or (self.multi_closure and len(solutions) > 1):
            solution = solutions[0]
        else:
            solution = NO_SOLUTION
        ok = solution != NO_SOLUTION and (len(loop) == 0 or not self.tr_safe)
        if ok and is_gap:
            if abs(len(solution[2]) - (bstop + 1 - bstart)) > self.gap_margin:
                ok = False
        if ok:
            return solution
        if is_gap or (self.fix_breaks and len(loop) == 0):
            from_right = self.consensus_from_right(right_paths)
            from_left = self.consensus_from_left(left_paths)
            new_start = start + len(from_left)
            new_stop = stop - len(from_right)
            if (new_start >= bstart + MIN_EXTEND or new_stop <= bstop - MIN_EXTEND) and \
                    not self.partial_matches_reference(start, from_left, from_right, stop, len(loop)):
                new_gap = max(self.min_gap, new_stop - new_start) if is_gap else self.min_gap
                seq = from_left + "N" * new_gap + from_right
                sol = self.trim_patch(start, seq, stop)
                return sol if sol[2] != "" else NO_SOLUTION
            return NO_SOLUTION
        return NO_SOLUTION

    @staticmethod
    def consensus_from_left(seqs):                    # GapFiller.scala:100-108
        s0 = seqs[0]
        if len(seqs) == 1:
            return s0
        m = min(len(s) for s in seqs)
        for i in range(m):
            for s in seqs[1:]:
                if s[i] != s0[i]:
                    return s0[:i]
        return s0[:m]

    @staticmethod
    def consensus_from_right(seqs):                   # GapFiller.scala:110-121
        s0 = seqs[0]
        if len(seqs) == 1:
            return s0
        m = min(len(s) for s in seqs)
        for i in range(m):
            for s in seqs[1:]:
                if s[len(s) - 1 - i] != s0[len(s0) - 1 - i]:
                    return s0[len(s0) - i:]
        return s0[len(s0) - m:]

    def assemble_into_break(self, bstart, bstop):     # GapFiller.scala:123-158
        asm = Assembler(self.graph, self.K, self.tr_safe)
        r = self.region
        start = max(bstart - self.break_radius, r.start)
        stop = min(bstop + 1 + self.break_radius, r.stop)
        left = r.sub_string(start, bstart - start)
        right = r.sub_string(bstop + 1, stop - bstop - 1)
        forward, reverse, loop = asm.multi_bridge(left, right)
        return start, forward, reverse, stop, loop

    def break_joins(self, start, forward_paths, reverse_paths, stop):   # GapFiller.scala:160-181
        found = []
        for f in forward_paths:
            for rv in reverse_paths:
                s = self.join_break(start, f, rv, stop)
                if s != NO_SOLUTION and s not in found:
                    found.append(s)
        solutions = sorted(found, key=lambda a: len(a[2]) + len(a[1]))
        lengths = {len(s[2]) - len(s[1]) for s in found}
        return [solutions[0]] if len(lengths) == 1 else solutions

    def join_break(self, start, forward, reverse, stop):                # GapFiller.scala:183-212
        patch = self.proper_overlap(forward, reverse, self.k)
        if patch != "":
            return self.trim_patch(start, patch, stop)
        return NO_SOLUTION

    @staticmethod
    def proper_overlap(left, right, min_overlap):                      # GapFiller.scala:214-237
        ll, rl = len(left), len(right)
        for overlap in range(min_overlap, ll + rl - 2 * min_overlap + 1):
            lo = max(ll - overlap, 0)
            ro = max(overlap - ll, 0)
            n = min(ll - lo, rl - ro)
            if n <= 0 or left[lo:lo + n] == right[ro:ro + n]:      # substrMatch: an empty loop matches
                return left[:lo] + right[ro:]
        return ""

    def trim_patch(self, start, patch, stop):                          # GapFiller.scala:239-253
        r = self.region
        while start < stop and patch and (r.base_at(start) == patch[0] or r.original_base_at(start) == patch[0]):
            start += 1
            patch = patch[1:]
        while start < stop and patch and (r.base_at(stop - 1) == patch[-1] or
                                          r.original_base_at(stop - 1) == patch[-1]):
            stop -= 1
            patch = patch[:-1]
        return (start, r.ref_sub_string(start, stop - start), patch)

    def partial_matches_reference(self, start, from_left, from_right, stop, loop_length):  # GapFiller.scala:255-262
        r = self.region
        if start + len(from_left) > r.stop or stop - len(from_right) < r.start:
            return False
        region_left = r.sub_string(start, len(from_left))
        region_right = r.sub_string(stop - len(from_right), len(from_right))
        return from_left == region_left and from_right == region_right


def assemble_break(cfg, reads, bases, contig, start, stop, bstart, bstop, is_gap, K=47, min_depth=5,
                   break_radius=None, tr_safe=True, multi_closure=False, fix_breaks=False,
                   gap_margin=100000, min_gap=10):
    """GapFiller.fillGap (is_gap) or fixBreak over a region [start, stop] with
    its current bases and the contig as read, given the recruited reads:
    (status, (fix start, ref, patch), tandem repeat); status 4 where Pilon
    throws."""
    from oracle import oracle as O
    st, graph = O.kmer_graph(cfg, reads, K, min_depth)
    if st:
        return st, NO_SOLUTION, ""
    g = GapFiller(Region(start, stop, bases, contig), graph, K, break_radius, tr_safe, multi_closure,
                  fix_breaks, gap_margin, min_gap)
    try:
        sol = g.assemble_across_break(bstart, bstop, is_gap)
    except Edge:
        return 4, NO_SOLUTION, ""
    return 0, sol, g.tandem_repeat


# ------------------------------------------------------------------ region
# identifyAndFixIssues' local reassembly (GenomeRegion.scala:307-413) over
# the C oracle's region run: its changes (the snps/indels fix lists), its
# Pilon.bed gap and break regions, and recruitment from the read sets.

CIGAR_OPS = "MIDNSHP=X"


def valid_read(cfg, f):                               # BamFile.validateRead (BamFile.scala:100-104)
    return (cfg.nonpf or not f & 0x200) and (cfg.duplicates or not f & 0x400) and not f & 0x100


def recruit_flank(cfg, batch, start, stop, flank):
    """BamFile.readsInRegion over flankRegion (BamFile.scala:339-364): the
    indices of the valid reads whose span [pos, pos + reference length - 1]
    (one base when unmapped or of no reference length) overlaps
    [max(start - flank, 1), stop + flank]."""
    fs, fe = max(start - flank, 1), stop + flank
    idx = []
    for i in range(batch.n):
        f = int(batch.flag[i])
        if not valid_read(cfg, f):
            continue
        ps = int(batch.pos[i])
        words = batch.cigar[int(batch.cigar_off[i]):int(batch.cigar_off[i + 1])]
        rl = sum(int(w) >> 4 for w in words if int(w) & 15 in (0, 2, 3, 7, 8))
        pe = ps if (f & 0x4 or rl == 0) else ps + rl - 1
        if ps <= fe and pe >= fs:
            idx.append(i)
    return idx


def read_dicts(batch, idx):
    """The reads at idx as the dicts pack_reads takes."""
    from pilon_amd.abi import PL_RF_MATE_SAME_REF, PL_RF_NO_QUAL
    from pilon_amd.reads import unpack_2bit, unpack_nib
    nibs = "=ACMGRSVTWYHKDBN"
    exo = {int(r): int(o) for r, o in zip(batch.exo_read, batch.exo_off)}
    out = []
    for i in idx:
        s0, s1 = int(batch.seq_off[i]), int(batch.seq_off[i + 1])
        if i in exo:
            seq = "".join(nibs[c] for c in unpack_nib(batch.nib, exo[i], s1 - s0))
        else:
            seq = "".join("ACGT"[c] for c in unpack_2bit(batch.bases2, s0, s1 - s0))
        f = int(batch.flag[i])
        words = batch.cigar[int(batch.cigar_off[i]):int(batch.cigar_off[i + 1])]
        cig = "".join("%d%s" % (int(w) >> 4, CIGAR_OPS[int(w) & 15]) for w in words) or "*"
        out.append(dict(pos=int(batch.pos[i]), tlen=int(batch.tlen[i]), flag=f & 0x0FFF, mapq=int(batch.mapq[i]),
                        cigar=cig, seq=seq or "*",
                        qual=None if f & PL_RF_NO_QUAL else [int(q) for q in batch.quals[s0:s1]],
                        mate_same_ref=bool(f & PL_RF_MATE_SAME_REF)))
    return out


def fix_fix_list(fixes):                              # GenomeRegion.fixFixList (GenomeRegion.scala:557-593)
    fixes = sorted(fixes, key=lambda f: f[0])         # sortWith is stable
    out = []
    while fixes:
        if len(fixes) == 1:
            out.append(fixes[0])
            break
        f1, f2 = fixes[0], fixes[1]
        e1 = f1[0] + max(len(f1[1]) - 1, 0)
        e2 = f2[0] + max(len(f2[1]) - 1, 0)
        if f2[0] <= e1 and e2 >= f1[0]:               # Region.overlaps
            keep = f1 if len(f1[1]) + len(f1[2]) >= len(f2[1]) + len(f2[2]) else f2
            fixes = [keep] + fixes[2:]
        else:
            out.append(f1)
            fixes = fixes[1:]
    return out


def fix_issues(bases, start, fixes):                  # GenomeRegion.fixIssues (GenomeRegion.scala:595-621)
    for loc, was, patch in reversed(fix_fix_list(fixes)):
        i = loc - start
        bases = bases[:i] + patch + bases[i + len(was):]
    return bases


def region_string(name, s, e):                        # Region.regionString (Region.scala:42)
    return "%s:%d" % (name, s) if e + 1 - s < 2 else "%s:%d-%d" % (name, s, e)


def write_changes(fixes, name, new_name, offset):     # GenomeRegion.writeChanges (GenomeRegion.scala:645-657)
    out, delta = [], 0
    for loc, was, patch in fix_fix_list(fixes):
        new_loc = loc + delta
        out.append("%s %s %s %s\n" % (region_string(name, loc, loc + len(was) - 1),
                                      region_string(new_name, new_loc + offset, new_loc + offset + len(patch) - 1),
                                      was or ".", patch or "."))
        delta += len(patch) - len(was)
    return "".join(out)


def log_fix(name, reg, sol, gap_size, tandem_repeat):  # GenomeRegion.logFix (GenomeRegion.scala:472-502)
    loc, ref, patch = sol
    n_ref, n_patch = ref.count("N"), patch.count("N")
    if loc == 0:
        msg = "NoSolution"
    elif gap_size == 0 and not ref and not patch:
        msg = "NoChange"
    elif gap_size == 0 and n_patch == 0:
        msg = "BreakFix"
    elif n_patch > 0 and n_ref == 0:
        msg = "OpenedGap"
    elif gap_size > 0 and n_ref == gap_size and n_patch == 0:
        msg = "ClosedGap"
    elif gap_size > 0:
        msg = "PartialFill"
    else:
        msg = "Unknown!"
    outcome = "TandemRepeat" if tandem_repeat else msg
    if tandem_repeat:
        msg += " TandemRepeat %d" % len(tandem_repeat)
    line = "fix %s: %s %d -%d +%d %s\n" % ("gap" if gap_size > 0 else "break", region_string(name, *reg), loc,
                                           len(ref) - n_ref, len(patch) - n_patch, msg)
    return outcome, line


def reassemble_region(cfg, contig, start, stop, sets, name="ctg", new_name=None, offset=0, break_radius=None,
                      K=47, min_depth=5, tr_safe=True, multi_closure=False, fix_breaks=False, gap_margin=100000,
                      min_gap=10, fix_gaps=True, fix_local=True, vcf=False):
    """One region through identifyAndFixIssues with its gap filling and
    break fixing (GenomeRegion.scala:307-413, --nostrays), then writeChanges,
    writeVcf when vcf (before Pilon.bed, whose change regions then include
    the loci the reassembly records replace) and Pilon.bed.  sets:
    [(ReadBatch, BAM type name, maxInsertSize)].  Returns a dict: status (4
    where Pilon throws), changes, bases, vcf, bed (the region's lines with
    the reassembly outcomes), log (logFix lines); changes_py and bases_py
    are writeChanges and fixIssues restated here over the same lists."""
    from oracle import oracle as O
    from pilon_amd.abi import BAM_TYPES
    from pilon_amd.reads import pack_reads
    if new_name is None:
        new_name = name + "_pilon"
    res = O.run_region(cfg, contig, start, stop, [b for b, _, _ in sets], name=name, new_name=new_name,
                       offset=offset, vcf=O.BED, set_types=[BAM_TYPES[t] for _, t, _ in sets])
    if res["status"]:
        return {"status": res["status"]}
    contig_s = contig.decode() if isinstance(contig, bytes) else contig
    # the snps/indels fix lists from the changes (Pass 1 changes never overlap)
    snp, small = [], []
    for line in res["changes"].splitlines():
        old, _, was, patch = line.split(" ")
        loc = int(old.rsplit(":", 1)[1].split("-")[0])
        was, patch = ("" if was == "." else was), ("" if patch == "." else patch)
        (snp if len(was) == 1 and len(patch) == 1 and patch in "ACGT" else small).append((loc, was, patch))
    region_bases = contig_s[start - 1:stop]
    bases = fix_issues(region_bases, start, snp)      # fixIssues(snpFixList) (:380)
    gaps, breaks = [], []
    for line in res["bed"].splitlines():
        f = line.split("\t")
        if len(f) > 3 and f[3] in ("Gap", "Break"):
            (gaps if f[3] == "Gap" else breaks).append((int(f[1]) + 1, int(f[2])))
    long_run = bool(cfg.longread)
    big, bed_extra, log = [], [], []
    attempts = ([(g, True) for g in gaps] if fix_gaps else []) + ([(b, False) for b in breaks] if fix_local else [])
    for reg, is_gap in attempts:
        recruited = []
        for b, t, m in sets:                           # recruitFrags
            if t == "frags":
                recruited += read_dicts(b, recruit_flank(cfg, b, reg[0], reg[1], m))
        if not long_run:                               # recruitUnpaired
            for b, t, m in sets:
                if t != "frags":
                    recruited += read_dicts(b, recruit_flank(cfg, b, reg[0], reg[1], m))
        st, sol, loop = assemble_break(cfg, pack_reads(recruited), bases, contig_s, start, stop, reg[0], reg[1],
                                       is_gap, K=K, min_depth=min_depth, break_radius=break_radius, tr_safe=tr_safe,
                                       multi_closure=multi_closure, fix_breaks=fix_breaks, gap_margin=gap_margin,
                                       min_gap=min_gap)
        if st:
            return {"status": st}
        logged = None
        if is_gap:
            if sol[0] > 0:
                big.append(sol)
                logged = log_fix(name, reg, sol, reg[1] + 1 - reg[0], loop)
        elif sol[0] > 0 and max(len(sol[1]), len(sol[2])) > 10:
            big.append(sol)
            logged = log_fix(name, reg, sol, 0, loop)
        elif sol[0] == 0:
            outcome, line = log_fix(name, reg, sol, 0, loop)
            logged = (outcome, "# " + line)
        if logged:
            bed_extra.append("%s\t%d\t%d\t%s\t0\t+\t%d\t%d\t0,128,0\n" % (name, reg[0] - 1, reg[1], logged[0],
                                                                        reg[0] - 1, reg[1]))
            log.append(logged[1])
    big_list = list(reversed(big))                     # bigFixList ::= fix
    final = fix_issues(bases, start, small + big_list)  # fixIssues(smallFixList ++ bigFixList)
    res2 = O.run_region(cfg, contig, start, stop, [b for b, _, _ in sets], name=name, new_name=new_name,
                        offset=offset, vcf=O.BED | (O.VCF_ON if vcf else 0),
                        set_types=[BAM_TYPES[t] for _, t, _ in sets], big=big_list)
    if res2["status"]:
        return {"status": res2["status"]}
    return {"status": 0, "changes": res2["changes"], "bases": res2["bases"].decode(), "vcf": res2.get("vcf"),
            "bed": res2["bed"] + "".join(bed_extra), "log": "".join(log), "run": res,
            "changes_py": write_changes(snp + small + big_list, name, new_name, offset), "bases_py": final}
