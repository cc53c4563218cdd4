"""Hand-derived known-answer cases for Pilon's pileup path.

Each case is a tiny contig + reads whose expected counters were worked out by
hand from the Scala semantics (file:line cited per case).  The expected values
are literals, not recomputed by any model, so a dropped term, a wrong offset
or a swapped operand in either the oracle or the GPU engine fails them.
Default config: flank 10, minMq 0, minQual 0, defaultQual 10, minMinDepth 5.
"""
from pilon_amd.reads import pack_reads

# 36 bp, no homopolymers beyond 'CCC' at 16-18 (used by the deletion case)
REF_A = "ACGTAGCTAGCTTGACCCATGCATGCAAGTTCAGCA"
# 40 bp with 'ACAC' at 13-16 for the insertion rotation case
REF_B = "GCTAGCTAGCTGACACTTGCATGGATCCTGACGTTGCATG"
# 40 bp for the two-base deletion call; positions 14/16 differ (no slide)
REF_C = "GATCGATTGCAGTCAGCTTACGGATCCAGTAGCATGCATC"


def q_ramp(n, base=20):
    return [base + i for i in range(n)]


def read(pos, cigar, seq, qual=None, flag=0, mapq=60, tlen=0, mate_same_ref=True):
    return dict(pos=pos, cigar=cigar, seq=seq, qual=qual, flag=flag, mapq=mapq, tlen=tlen,
                mate_same_ref=mate_same_ref)


def case_trusted_flank():
    """PileUpRegion.scala:118,184-193; PileUp.scala:75-84; physCovIncr :62-88.
    One unpaired 30M read at 1, quals 30, MAPQ 60: only offsets 10..19 are
    trusted -> loci 11..20 get count 1, qualSum 30*61, mqSum 61, qSum 30.
    Unpaired fragment [1, 30): physCov 1 at 1..29, insertSize 29."""
    reads = [read(1, "30M", REF_A[0:30], [30] * 30)]
    exp = {}
    for locus in range(11, 21):
        b = "ACGT".index(REF_A[locus - 1])
        exp[locus] = {f"count_{'ACGT'[b]}": 1, f"qual_{'ACGT'[b]}": 1830, "mqSum": 61, "qSum": 30,
                      "physCov": 1, "insertSize": 29}
    exp[1] = {"physCov": 1, "insertSize": 29, "mqSum": 0}
    exp[29] = {"physCov": 1, "insertSize": 29, "mqSum": 0}
    exp[30] = {"physCov": 0, "insertSize": 0}
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect=exp, stats={"base_count": 10, "read_count": 1})


def case_soft_clip():
    """PileUpRegion.scala:139-141,194-206.  5S25M at 11: clipStart 6, clipEnd
    10 (clips +1 each), badPair +1 at 6..10; adjMq = roundDiv(60*25, 30) = 50,
    so mq1 51 and qualSum 30*51 = 1530 at the trusted loci 16..25."""
    seq = "TTTTT" + REF_A[10:35]
    reads = [read(11, "5S25M", seq, [30] * 30)]
    exp = {6: {"clips": 1, "badPair": 1}, 7: {"clips": 0, "badPair": 1}, 10: {"clips": 1, "badPair": 1},
           11: {"badPair": 0, "mqSum": 0}, 15: {"mqSum": 0}}
    for locus in range(16, 26):
        b = REF_A[locus - 1]
        exp[locus] = {f"count_{b}": 1, f"qual_{b}": 1530, "mqSum": 51, "qSum": 30}
    exp[26] = {"mqSum": 0}
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect=exp, stats={"base_count": 10})


def case_invalid_pair():
    """PileUpRegion.scala:40-46,107: a paired read without the proper flag is
    invalid: its trusted bases go to badPair only, no counts, no baseCount,
    no physical coverage."""
    reads = [read(1, "30M", REF_A[0:30], [30] * 30, flag=0x1 | 0x40, tlen=200)]
    exp = {locus: {"badPair": 1, "mqSum": 0, "qSum": 0, "physCov": 0} for locus in range(11, 21)}
    exp[10] = {"badPair": 0}
    exp[21] = {"badPair": 0}
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect=exp, stats={"base_count": 0, "read_count": 1})


def case_deletion_slide():
    """PileUpRegion.scala:163-183.  17M1D13M at 1 deletes the third C of CCC
    (locus 18).  It slides to 16; on each step the slid-over read base is
    added len bases later (remove() is a no-op for valid reads): the C of
    offset 16 (q36) lands on 18 and the C of offset 15 (q35) on 17, so 17 and
    18 count that base twice.  The deletion is recorded at 16 with the
    pre-normalisation quality quals(17) = 37; it adds mq1 to mqSum."""
    seq = REF_A[0:17] + REF_A[18:31]
    reads = [read(1, "17M1D13M", seq, q_ramp(30))]
    exp = {
        15: {"count_A": 1, "qual_A": 34 * 61, "mqSum": 61, "qSum": 34},
        16: {"count_C": 1, "qual_C": 35 * 61, "mqSum": 122, "qSum": 35 + 37, "deletions": 1, "delQual": 61},
        17: {"count_C": 2, "qual_C": (36 + 35) * 61, "mqSum": 122, "qSum": 71},
        18: {"count_C": 1, "qual_C": 36 * 61, "mqSum": 61, "qSum": 36, "deletions": 0},
        19: {"count_A": 1, "qual_A": 37 * 61, "qSum": 37},
        21: {"count_G": 1, "qual_G": 39 * 61, "qSum": 39},
        22: {"mqSum": 0},
    }
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect=exp, stats={"base_count": 12})


def case_insertion_rotation(n_reads=6, start=1):
    """PileUpRegion.scala:150-162; PileUp.scala:98-105,209-246.
    16M2I12M with inserted AC after locus 16 of ...G ACAC T...: the insertion
    slides left while the base before it equals its last base, rotating
    AC->CA->AC->CA->AC, and stops at 13 (G before it).  Six reads: insertions
    6, insQual 366, qSum at 13 = 6*(32+36) (base q32 + insertion q36); mqSum
    only counts the bases (insertion adds no mqSum).  insPct 100 > high 65
    (L=2: middle 43, low 21) -> homozygous insertion AC at 13, fixed."""
    seq = REF_B[0:16] + "AC" + REF_B[16:28]
    reads = [read(1, "16M2I12M", seq, q_ramp(30)) for _ in range(n_reads)]
    exp = {13: {"count_A": n_reads, "qual_A": n_reads * 32 * 61, "mqSum": n_reads * 61,
                "qSum": n_reads * (32 + 36), "insertions": n_reads, "insQual": n_reads * 61},
           17: {"insertions": 0, "count_T": n_reads, "qSum": n_reads * 38}}
    fixed = REF_B[0:12] + "AC" + REF_B[12:]
    changes = "ctg:13 ctg_pilon:13-14 . AC\n"
    return dict(ref=REF_B, start=start, stop=40, reads=reads, expect=exp,
                changes=changes if start == 1 else None, bases=fixed if start == 1 else None,
                stats={"base_count": n_reads * 8, "ins": 1, "ins_bases": 2, "min_depth": 5})


def case_deletion_call():
    """GenomeRegion.scala:255-264; PileUp.scala:188-191,209-246.  Six reads
    14M2D16M delete loci 15-16 (no slide: locus 14 'C' != locus 16 'T').  At
    15: count 0, deletions 6, delPct 100 > 65 -> homozygous DEL 'AG'; Pass 1
    marks 16 deleted and adds pu.deletions (6) into its pileup."""
    seq = REF_C[0:14] + REF_C[16:32]
    reads = [read(1, "14M2D16M", seq, [30] * 30) for _ in range(6)]
    exp = {15: {"deletions": 6, "delQual": 366, "mqSum": 366, "qSum": 180, "count_A": 0},
           16: {"deletions": 6, "delQual": 0, "mqSum": 0}}
    fixed = REF_C[0:14] + REF_C[16:]
    changes = "ctg:15-16 ctg_pilon:15 %s .\n" % REF_C[14:16]
    return dict(ref=REF_C, start=1, stop=40, reads=reads, expect=exp, changes=changes, bases=fixed,
                stats={"base_count": 60, "dels": 1, "del_bases": 2},
                deleted=[16], disp={15: 4, 16: 0})


def case_deletion_n_anchor():
    """Vcf.scala:84-91,131,138-141: case_deletion_call with an N in the draft
    at 14, the base before the deletion.  Pass 1 still calls the DEL at 15
    (its own reference base is A, GenomeRegion.scala:255); 14 itself gets no
    disposition (r == 'N').  The VCF prints the deletion at 14 as REF "NAG"
    with cBase "N", so ALT is "." and AF 0.00."""
    ref = REF_C[:13] + "N" + REF_C[14:]
    seq = REF_C[0:14] + REF_C[16:32]
    reads = [read(1, "14M2D16M", seq, [30] * 30) for _ in range(6)]
    exp = {14: {"count_C": 6}, 15: {"deletions": 6, "delQual": 366, "mqSum": 366, "qSum": 180, "count_A": 0},
           16: {"deletions": 6, "delQual": 0, "mqSum": 0}}
    return dict(ref=ref, start=1, stop=40, reads=reads, expect=exp, changes="ctg:15-16 ctg_pilon:15 AG .\n",
                bases=ref[0:14] + ref[16:], stats={"base_count": 60, "dels": 1, "del_bases": 2},
                deleted=[16], disp={14: 0, 15: 4, 16: 0})


def case_physcov_pairs():
    """PileUpRegion.scala:62-100.  Region 101-200 of a 300 bp contig.  A proper
    pair starting at 50 with TLEN 120 covers [50,170): it enters through
    physCovStart.  A second starting at 150 with TLEN 100 covers [150,250).
    Their mates (TLEN < 0) add nothing.  physCov 1/2/1 and insertSize
    120, (120+100)/2 = 110, 100 on 101-149 / 150-169 / 170-200."""
    ref = (REF_A * 9)[:300]
    pf = 0x1 | 0x2
    reads = [read(50, "30M", ref[49:79], [30] * 30, flag=pf | 0x40 | 0x20, tlen=120),
             read(140, "30M", ref[139:169], [30] * 30, flag=pf | 0x80 | 0x10, tlen=-120),
             read(150, "30M", ref[149:179], [30] * 30, flag=pf | 0x40 | 0x20, tlen=100),
             read(220, "30M", ref[219:249], [30] * 30, flag=pf | 0x80 | 0x10, tlen=-100)]
    exp = {101: {"physCov": 1, "insertSize": 120}, 149: {"physCov": 1, "insertSize": 120},
           150: {"physCov": 2, "insertSize": 110}, 169: {"physCov": 2, "insertSize": 110},
           170: {"physCov": 1, "insertSize": 100}, 200: {"physCov": 1, "insertSize": 100}}
    return dict(ref=ref, start=101, stop=200, reads=reads, expect=exp, stats={"read_count": 4})


def case_n_base():
    """PileUp.scala:46-52,77 and PileUpRegion.scala:40-43: an N at a trusted
    offset is counted in the region baseCount but not in the pileup."""
    seq = REF_A[0:12] + "N" + REF_A[13:30]
    reads = [read(1, "30M", seq, [30] * 30)]
    exp = {13: {"mqSum": 0, "qSum": 0, "count_A": 0, "count_C": 0, "count_G": 0, "count_T": 0},
           12: {"mqSum": 61}}
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect=exp, stats={"base_count": 10})


def case_no_qual():
    """PileUpRegion.scala:114-115: a read without QUAL uses defaultQual 10."""
    reads = [read(1, "30M", REF_A[0:30], None)]
    exp = {11: {"qSum": 10, "mqSum": 61}}
    b = REF_A[10]
    exp[11][f"qual_{b}"] = 610
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect=exp)


def case_ambiguous(iupac=False):
    """GenomeRegion.scala:255-271,326-341; PileUp.scala:132-167; Bases.scala:67-90.
    Ten 30M reads at 5 (trusted loci 15..24) carry C (five) or G (five) at
    locus 20 (ref T), q30, MAPQ 60.  qualSum C = G = 5*30*61 = 9150; the
    stable order keeps C before G (base C, alt G); total 18300, homoScore 0,
    heteroScore 18300 -> not homo; score = 18300*10/610 = 300, q 30.
    b != r, score > 0, not homo, alt != r -> AMB.  The fix is (20, T, C); with
    --iupac it goes to the small fix list as toIUPAC(C, G) = S."""
    reads = []
    for k in range(10):
        seq = REF_A[4:19] + ("C" if k < 5 else "G") + REF_A[20:34]
        reads.append(read(5, "30M", seq, [30] * 30))
    base = "S" if iupac else "C"
    exp = {20: {"count_C": 5, "count_G": 5, "qual_C": 9150, "qual_G": 9150, "mqSum": 610, "qSum": 300}}
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect=exp, changes=f"ctg:20 ctg_pilon:20 T {base}\n",
                bases=REF_A[:19] + base + REF_A[20:], disp={20: 5}, stats={"amb": 1, "snps": 0},
                cfg={"iupac": 1} if iupac else {})


def case_ambiguous_iupac():
    return case_ambiguous(iupac=True)


def case_het_deletion(old_indel=False):
    """PileUp.scala:209-246; GenomeRegion.scala:255-271.  Six reads 14M2D16M
    delete loci 15-16 (as in case_deletion_call) beside twelve 30M reads of
    the reference.  At 15: count 12 (A), deletions 6, mqSum 12*61 + 6*61 =
    1098, delQual 366, qSum 12*30 + 6*30 = 540; delPct = max(roundDiv(36600,
    1098), roundDiv(600, 18)) = 33.  New heuristics (L=2: middle 43, low 21,
    high 65): het deletion -> not homoIndel, no DEL, not confirmed (indel) ->
    disposition NONE.  --oldindel: 33 >= 33 but 33 < 50-2 -> no indel call;
    homo A, q = 21960*12/1098/12 = 20 -> CONFIRMED.  No change either way."""
    seq = REF_C[0:14] + REF_C[16:32]
    reads = [read(1, "14M2D16M", seq, [30] * 30) for _ in range(6)]
    reads += [read(1, "30M", REF_C[0:30], [30] * 30) for _ in range(12)]
    exp = {15: {"count_A": 12, "deletions": 6, "delQual": 366, "mqSum": 1098, "qSum": 540}}
    return dict(ref=REF_C, start=1, stop=40, reads=reads, expect=exp, changes="", bases=REF_C,
                disp={15: 1 if old_indel else 0}, stats={"dels": 0},
                cfg={"old_indel": 1} if old_indel else {})


def case_het_deletion_oldindel():
    return case_het_deletion(old_indel=True)


def case_min_depth_300(min_depth=300.0):
    """GenomeRegion.scala:221-224,255-256: --mindepth 300 (>= 1: taken as an
    Int).  310 reads 30M at 1, q30 MAPQ 60: loci 11..20 hold 310 bases
    bases each, homozygous on the reference at q = 30 -> confirmed when
    310 >= minDepth: 10 confirmed at 300, none at 311."""
    reads = [read(1, "30M", REF_A[0:30], [30] * 30) for _ in range(310)]
    n = 10 if min_depth <= 310 else 0
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect={11: {"count_C": 310}},
                stats={"min_depth": int(min_depth), "n_confirmed": n, "base_count": 3100},
                disp={11: 1 if n else 0, 21: 0}, cfg={"min_depth": float(min_depth)})


def case_min_depth_311():
    return case_min_depth_300(311.0)


def case_deep_region():
    """GenomeRegion.scala:221-224,255-256 at the default minDepth 0.1: 6000
    reads 36M at 1 (q30, MAPQ 60) pile 16 trusted bases each at loci 11..26;
    baseCount 96000 over 36 loci -> coverage roundDiv = 2667, minDepth =
    round(266.7) = 267 (above a 256-bucket depth histogram).  Loci 11..26
    have depth 6000, homozygous reference, q 30 -> 16 confirmed."""
    reads = [read(1, "36M", REF_A, [30] * 36) for _ in range(6000)]
    return dict(ref=REF_A, start=1, stop=36, reads=reads, expect={11: {"count_C": 6000, "qual_C": 6000 * 30 * 61}},
                stats={"min_depth": 267, "coverage": 2667, "n_confirmed": 16, "base_count": 96000},
                disp={11: 1, 26: 1, 27: 0})


def case_deep_70k():
    """BaseSum.scala:24-30, PileUp.scala:75-84: 70,000 reads 36M at 1 (q40,
    MAPQ 60) pile 70,000 bases at each of loci 11..26 -- past 65,536, with
    qualSum 70,000 * 40 * 61 = 170,800,000 (still an Int).  baseCount
    1,120,000 / 36 -> coverage 31,111, minDepth round(3111.1) = 3111; the 16
    loci are confirmed."""
    reads = [read(1, "36M", REF_A, [40] * 36) for _ in range(70000)]
    return dict(ref=REF_A, start=1, stop=36, reads=reads,
                expect={11: {"count_C": 70000, "qual_C": 170800000, "mqSum": 4270000, "qSum": 2800000}},
                stats={"min_depth": 3111, "coverage": 31111, "n_confirmed": 16, "base_count": 1120000},
                disp={11: 1, 26: 1})


WIDE_QUAL = 2_304_000_000   # 90,000 * 100 * 256


def case_wide_qualsum():
    """BaseSum.scala:24 (Long sums), PileUp.scala:132-167: 90,000 reads 36M
    at 1 with q100 and MAPQ 255 (mq1 256) give qualSum 90,000 * 100 * 256 =
    2,304,000,000 at loci 11..26, beyond an Int: the pileup plane holds its
    low 32 bits (-1,990,967,296) and the call is made from the Long:
    homoScore 2.304e9, heteroScore 0, score = 2.304e9 * 90,000 / mqSum
    (23,040,000) = 9,000,000, q = 100, confirmed.  mqSum 23,040,000 and
    qSum 9,000,000 stay Ints."""
    reads = [read(1, "36M", REF_A, [100] * 36, mapq=255) for _ in range(90000)]
    low = WIDE_QUAL - 2 ** 32
    return dict(ref=REF_A, start=1, stop=36, reads=reads,
                expect={11: {"count_C": 90000, "qual_C": low, "mqSum": 23040000, "qSum": 9000000}},
                stats={"n_confirmed": 16, "base_count": 1440000}, disp={11: 1, 26: 1}, wide=list(range(11, 27)))


ALL = [case_trusted_flank, case_soft_clip, case_invalid_pair, case_deletion_slide, case_insertion_rotation,
       case_deletion_call, case_deletion_n_anchor, case_physcov_pairs, case_n_base, case_no_qual, case_ambiguous, case_ambiguous_iupac,
       case_het_deletion, case_het_deletion_oldindel, case_min_depth_300, case_min_depth_311, case_deep_region,
       case_deep_70k, case_wide_qualsum]

# ---- long-read BAMs (--nanopore / --pacbio: PileUpRegion.addRead with
# longRead > 0, GenomeRegion.excludeMotifs).  A case's "sets" lists
# (reads, BAM type) pairs; the other cases are one frags BAM.
LR = {"longread": 1}
# 40 bp: AAAAA at loci 13-17 / AAAA at 13-16, G at 12, T after the run
REF_H5 = "GCTAGCTAGCTG" + "AAAAA" + "TCGATCGTACGTTGCATGCATGC"
REF_H4 = "GCTAGCTAGCTG" + "AAAA" + "TCGATCGTACGTTGCATGCATGCA"
# 43 bp: CCAGG at loci 20-24 and no other CC_GG
REF_N = "GATCGATTGCAGTCAGTTA" + "CCAGG" + "TATGCATCGATTCGATCAG"
# 42 bp: TTTT at loci 20-23 (G before, C after)
REF_T = "GATCGATTGCAGTCAGCTG" + "TTTT" + "CAGTAGCATGCATCGAGCG"


def case_lr_insertion_pacbio():
    """PileUpRegion.scala:142,160: case_insertion_rotation from a pacbio BAM.
    The insertion stops at 13, where homoRun(13) (refBases from index 13:
    C) is 1, so it is kept with indelMq = adjMq min 8 = 8: insQual 6*9.
    insPct is the larger of pct(insQual, mqSum) and pct(insertions, count)
    (PileUp.scala:122), 100, so the call and its fix stand; excludeMotifs
    (GenomeRegion.scala:219, every BAM long-read) leaves 13 alone: homoRun(13)
    over the region is 1 (A then C)."""
    c = case_insertion_rotation()
    c["expect"] = {13: {"insertions": 6, "insQual": 54, "qSum": 6 * (32 + 36), "mqSum": 6 * 61}}
    c["sets"] = [(c["reads"], "pacbio")]
    c["cfg"] = LR
    return c


def _h_reads(ref, run):
    # one extra A after the run: nM 1I (30 - n)M at 1, quals 30
    seq = ref[0:12 + run] + "A" + ref[12 + run:29]
    return [read(1, "%dM1I%dM" % (12 + run, 29 - 12 - run), seq, [30] * 30) for _ in range(6)]


def case_lr_insertion_homopolymer5():
    """PileUpRegion.scala:150-162: an extra A after AAAAA (13-17) slides to 13
    (locus 12 is G).  homoRun(iloc) reads refBases from 0-based index 13,
    i.e. loci 14-17, AAAA: 4, so a pacbio read's insertion is dropped (no
    insertion, no insQual, no insertion qual in qSum at 13)."""
    reads = _h_reads(REF_H5, 5)
    return dict(ref=REF_H5, start=1, stop=40, reads=reads, sets=[(reads, "pacbio")], cfg=LR,
                expect={13: {"insertions": 0, "insQual": 0, "qSum": 180, "count_A": 6}},
                stats={"ins": 0})


def case_lr_insertion_homopolymer5_frags():
    """The same reads from a frags BAM: the insertion is added at 13 (6 x
    insQual 61, qSum 30 + 30 per read) and fixed."""
    reads = _h_reads(REF_H5, 5)
    return dict(ref=REF_H5, start=1, stop=40, reads=reads,
                expect={13: {"insertions": 6, "insQual": 366, "qSum": 360}}, stats={"ins": 1, "ins_bases": 1})


def case_lr_insertion_homopolymer4():
    """An extra A after AAAA (13-16) also stops at 13, but homoRun(13) sees
    loci 14-16 then T: 3, so the pacbio insertion is kept (insQual 6*9).
    The INS call at 13 is not fixed: excludeMotifs' homoRun(13) over the
    region counts AAAA, 4 (GenomeRegion.scala:416-422,438)."""
    reads = _h_reads(REF_H4, 4)
    return dict(ref=REF_H4, start=1, stop=40, reads=reads, sets=[(reads, "pacbio")], cfg=LR,
                expect={13: {"insertions": 6, "insQual": 54, "qSum": 360}},
                stats={"ins": 0}, changes="", bases=REF_H4)


def _del_reads():
    # 15M1D15M at 1 deletes locus 16, the last A of AAAA (13-16)
    seq = REF_H4[0:15] + REF_H4[16:31]
    return [read(1, "15M1D15M", seq, [30] * 30) for _ in range(6)]


def case_lr_deletion_homopolymer():
    """PileUpRegion.scala:163-183: the deletion slides from 16 to 13, adding
    the slid-over A of each step one locus later (14, 15, 16 gain a count).
    homoRun(index(13)) = homoRun(12) reads loci 13-16, AAAA: 4, so the pacbio
    deletion is dropped: no deletions, delQual or deletion mqSum at 13."""
    reads = _del_reads()
    return dict(ref=REF_H4, start=1, stop=40, reads=reads, sets=[(reads, "pacbio")], cfg=LR,
                expect={13: {"deletions": 0, "delQual": 0, "mqSum": 6 * 61, "count_A": 6},
                        14: {"count_A": 12}, 15: {"count_A": 12}, 16: {"count_A": 6}},
                stats={"dels": 0})


def case_lr_deletion_homopolymer_region5():
    """The same reads in region 5-40: index(13) is 8, and homoRun(8) reads
    refBases(8) = G, then C: 1, so the deletion is kept at 13 with indelMq
    8 (delQual 6*9, mqSum 6*61 + 6*9).  Its DEL call is not fixed:
    excludeMotifs' homoRun(13) over the region is 4."""
    reads = _del_reads()
    return dict(ref=REF_H4, start=5, stop=40, reads=reads, sets=[(reads, "pacbio")], cfg=LR,
                expect={13: {"deletions": 6, "delQual": 54, "mqSum": 6 * 70, "qSum": 6 * 60}},
                stats={"dels": 0})


def _n_reads():
    return [read(2, "40M", REF_N[1:41], [30] * 40) for _ in range(6)]


def case_lr_nanopore_qual():
    """PileUpRegion.scala:128-134,190: a nanopore base at a locus whose
    index(locus) passes nanoporeExclude piles with quality 0.  Region 1-43:
    index(22) = 21, refBases 19,20 = CC and 22,23 = GG, so the A at 22
    counts with no quality (qual_A 0, qSum 0, mqSum 6*61)."""
    reads = _n_reads()
    return dict(ref=REF_N, start=1, stop=43, reads=reads, sets=[(reads, "nanopore")], cfg=LR,
                expect={22: {"count_A": 6, "qual_A": 0, "qSum": 0, "mqSum": 366},
                        21: {"count_C": 6, "qual_C": 6 * 30 * 61, "qSum": 180},
                        24: {"count_G": 6, "qual_G": 6 * 30 * 61}})


def case_lr_nanopore_qual_region3():
    """Region 3-43: the region index is read as a contig index, so the motif
    test passes for index 21 = locus 24 (refBases 19..23 = CCAGG): the G at
    24 piles with quality 0 and the A at 22 keeps its own."""
    reads = _n_reads()
    return dict(ref=REF_N, start=3, stop=43, reads=reads, sets=[(reads, "nanopore")], cfg=LR,
                expect={22: {"count_A": 6, "qual_A": 6 * 30 * 61, "qSum": 180},
                        24: {"count_G": 6, "qual_G": 0, "qSum": 0, "mqSum": 366}})


def _n_del_reads():
    # 21M1D19M at 1 deletes the A of CCAGG (locus 22); C at 21 != A: no slide
    seq = REF_N[0:21] + REF_N[22:41]
    return [read(1, "21M1D19M", seq, [30] * 40) for _ in range(6)]


def case_lr_nanopore_deletion():
    """PileUpRegion.scala:180-182: homoRun(index(22)) is 1 but
    nanoporeExclude(21) holds, so a nanopore deletion at 22 is dropped."""
    reads = _n_del_reads()
    return dict(ref=REF_N, start=1, stop=43, reads=reads, sets=[(reads, "nanopore")], cfg=LR,
                expect={22: {"deletions": 0, "delQual": 0, "mqSum": 0, "qSum": 0}})


def case_lr_pacbio_deletion():
    """The same deletion from a pacbio BAM is kept with indelMq 8: delQual
    and mqSum 6*9, qSum 6*30."""
    reads = _n_del_reads()
    return dict(ref=REF_N, start=1, stop=43, reads=reads, sets=[(reads, "pacbio")], cfg=LR,
                expect={22: {"deletions": 6, "delQual": 54, "mqSum": 54, "qSum": 180}})


def _snp_reads(ref, locus, base):
    seq = ref[0:locus - 1] + base + ref[locus:40]
    return [read(1, "40M", seq, [30] * 40) for _ in range(6)]


def case_lr_snp_homopolymer_excluded():
    """GenomeRegion.scala:219,326,432-439: with only long-read BAMs, a SNP at
    20, the first T of TTTT, is called but takes no fix and no count."""
    reads = _snp_reads(REF_T, 20, "C")
    return dict(ref=REF_T, start=1, stop=42, reads=reads, sets=[(reads, "pacbio")], cfg=LR,
                expect={20: {"count_C": 6}}, stats={"snps": 0}, changes="", bases=REF_T, disp={20: 2})


def case_lr_snp_homopolymer_with_frags():
    """The same with a frags BAM beside the pacbio one (fragBams non-empty,
    GenomeRegion.scala:147): the SNP is fixed."""
    reads = _snp_reads(REF_T, 20, "C")
    return dict(ref=REF_T, start=1, stop=42, reads=reads, sets=[(reads, "pacbio"), ([], "frags")], cfg=LR,
                expect={20: {"count_C": 6}}, stats={"snps": 1}, changes="ctg:20 ctg_pilon:20 T C\n",
                bases=REF_T[:19] + "C" + REF_T[20:])


def case_lr_snp_nanopore_motif():
    """A SNP at 22, the middle of CCAGG, from a pacbio BAM beside an empty
    nanopore BAM: excludeMotifs' nanoporeExclude(22) holds, so no fix."""
    reads = _snp_reads(REF_N, 22, "T")
    return dict(ref=REF_N, start=1, stop=43, reads=reads, sets=[(reads, "pacbio"), ([], "nanopore")], cfg=LR,
                expect={22: {"count_T": 6}}, stats={"snps": 0}, changes="", bases=REF_N)


def case_lr_snp_pacbio_motif():
    """The same pacbio reads alone: nothing excludes 22, the SNP is fixed."""
    reads = _snp_reads(REF_N, 22, "T")
    return dict(ref=REF_N, start=1, stop=43, reads=reads, sets=[(reads, "pacbio")], cfg=LR,
                expect={22: {"count_T": 6}}, stats={"snps": 1}, changes="ctg:22 ctg_pilon:22 A T\n",
                bases=REF_N[:21] + "T" + REF_N[22:])


LONG = [case_lr_insertion_pacbio, case_lr_insertion_homopolymer5, case_lr_insertion_homopolymer5_frags,
        case_lr_insertion_homopolymer4, case_lr_deletion_homopolymer, case_lr_deletion_homopolymer_region5,
        case_lr_nanopore_qual, case_lr_nanopore_qual_region3, case_lr_nanopore_deletion, case_lr_pacbio_deletion,
        case_lr_snp_homopolymer_excluded, case_lr_snp_homopolymer_with_frags, case_lr_snp_nanopore_motif,
        case_lr_snp_pacbio_motif]


_batches = {}


def batch_of(case):
    """The case's reads packed (cached per case: the deep cases hold tens of
    thousands of identical reads)."""
    key = (case["ref"], case["start"], case["stop"], len(case["reads"]), repr(case["reads"][:2]),
           repr(case["reads"][-1:]))
    if key not in _batches:
        _batches[key] = pack_reads(case["reads"])
    return _batches[key]


def sets_of(case):
    """[(batch, BAM type name)] of the case: its "sets", else one frags BAM."""
    if "sets" not in case:
        return [(batch_of(case), "frags")]
    return [(pack_reads(reads), t) for reads, t in case["sets"]]


LR_TYPE = {"nanopore": 1, "pacbio": 2}


def merged_of(case):
    """The case's BAMs as one coordinate-sorted batch (stable, in set order),
    with each read's longReadType and the set's kinds (bit 0: a frags BAM,
    bit 1: a nanopore BAM), as the engine merges a region's BAMs."""
    sets = case.get("sets", [(case["reads"], "frags")])
    tagged = [(r, LR_TYPE.get(t, 0)) for reads, t in sets for r in reads]
    tagged.sort(key=lambda x: x[0]["pos"])
    kinds = (1 if any(t == "frags" for _, t in sets) else 0) | (2 if any(t == "nanopore" for _, t in sets) else 0)
    return pack_reads([r for r, _ in tagged], sort=False), [t for _, t in tagged], kinds
