"""The oracle's region-level local reassembly (identifyAndFixIssues with
fillGap over the gaps and fixBreak over the possible breaks, --nostrays)
against what a planted draft fixes: SNPs, a run of N and a deletion made
from a known truth, which polishing must give back; and its fix-list
helpers on hand cases and against the C oracle's own writeChanges."""
import pytest

from pilon_amd.abi import default_config

from . import asm_cases as AC


def asm():
    from oracle import asm as A
    return A


def cfg(**kw):
    return default_config(fix_snps=1, fix_indels=1, **kw)


def run(c, **kw):
    d = c["draft"]
    return asm().reassemble_region(cfg(), d, 1, len(d), [(c["reads"], "frags", 500)], name="ctg", **kw)


@pytest.fixture(scope="module")
def planted():
    return AC.planted_region()


def test_planted_region_polishes_to_the_truth(planted):
    c = planted
    r = run(c)
    assert r["status"] == 0 and r["bases"] == c["truth"]
    g0, g1 = c["gap"]
    j = c["junction"]
    t = c["truth"]
    assert r["log"].splitlines()[0] == "fix gap: ctg:%d-%d %d -0 +100 ClosedGap" % (g0, g1, g0)
    brk = r["log"].splitlines()[1].split()
    assert brk[:2] == ["fix", "break:"] and brk[3:] == [str(j), "-0", "+40", "BreakFix"]
    lines = r["changes"].splitlines()
    assert "ctg:%d-%d ctg_pilon:%d-%d %s %s" % (g0, g1, g0, g1, "N" * 100, t[g0 - 1:g1]) in lines
    assert "ctg:%d ctg_pilon:%d-%d . %s" % (j, j, j + 39, t[j - 1:j + 39]) in lines
    assert sum(1 for l in lines if l.split()[2] != "." and len(l.split()[2]) == 1) == 3     # the SNPs
    outcomes = [l.split("\t")[3] for l in r["bed"].splitlines() if l.endswith("0,128,0")]
    assert outcomes == ["ClosedGap", "BreakFix"]


def test_each_fix_kind_can_be_left_out(planted):
    """--fix without gaps keeps the N run; without local the deletion stays."""
    c = planted
    t = c["truth"]
    g0, g1 = c["gap"]
    j = c["junction"]
    r = run(c, fix_gaps=False)
    assert r["bases"] == t[:g0 - 1] + "N" * 100 + t[g1:]
    r = run(c, fix_local=False)
    assert r["bases"] == t[:j - 1] + t[j + 39:]
    r = run(c, fix_gaps=False, fix_local=False)
    assert r["log"] == "" and r["changes"] == r["run"]["changes"]


def test_python_fix_lists_agree_with_the_c_oracle(planted):
    """The C oracle's fixIssues(smallFixList ++ bigFixList) and writeChanges
    over the big fixes given to it equal the Python restatement's."""
    r = run(planted)
    assert r["changes"] == r["changes_py"] and r["bases"] == r["bases_py"]
    r = run(AC.planted_region(seed=33, n=20000, gap=(4000, 120), deletion=(15000, 25), snps=(1000, 18000)))
    assert r["changes"] == r["changes_py"] and r["bases"] == r["bases_py"]


def test_fix_records_in_the_vcf(planted):
    """writeVcf with the reassembly's fixes: a writeFixRecord line at each
    big fix's locus (SVTYPE INS for the closed gap, whose was and patch are
    both 100 long, and for the restored deletion), the gap's loci embedded
    ("Del" filter, DP from the base count); the change regions of Pilon.bed
    then cover the replaced N run."""
    c = planted
    r = run(c, vcf=True)
    g0, g1 = c["gap"]
    j = c["junction"]
    t, d = c["truth"], c["draft"]
    lines = r["vcf"].splitlines()
    fixes = [l for l in lines if "SVTYPE=" in l and "<DUP>" not in l]
    assert fixes == [
        "ctg\t%d\t.\t%s\t%s\t.\tPASS\tSVTYPE=INS;SVLEN=0;END=%d\tGT\t1/1" % (g0 - 1, d[g0 - 2] + "N" * 100,
                                                                          d[g0 - 2] + t[g0 - 1:g1], g1),
        "ctg\t%d\t.\t%s\t%s\t.\tPASS\tSVTYPE=INS;SVLEN=40;END=%d\tGT\t1/1" % (j - 1, d[j - 2],
                                                                           d[j - 2] + t[j - 1:j + 39], j - 1)]
    at = lines.index(fixes[0])
    assert lines[at + 1].split("\t")[1] == str(g0)        # before the records of the fix's locus
    for k in range(100):
        assert lines[at + 1 + k].split("\t")[6] == "Del"
    assert lines[at + 101].split("\t")[6] != "Del"
    xs = [l.split("\t")[1:3] for l in r["bed"].splitlines() if l.split("\t")[3] == "X"]
    assert [str(g0 - 1), str(g1)] in xs
    assert [str(g0 - 1), str(g1)] not in [l.split("\t")[1:3] for l in run(c)["bed"].splitlines()
                                          if l.split("\t")[3] == "X"]


def test_write_changes_agrees_with_the_c_oracle():
    """With no reassembly the Python writeChanges over the recovered fix
    lists is the C oracle's own .changes text, offset and names included."""
    c = AC.planted_region(seed=32)
    d = c["draft"]
    r = asm().reassemble_region(cfg(), d, 1, len(d), [(c["reads"], "frags", 500)], name="ctg", new_name="x|y",
                                offset=1234, fix_gaps=False, fix_local=False)
    assert r["changes"] == r["run"]["changes"] and r["changes"].count("\n") == 3


def test_fix_fix_list_hand_cases():
    """GenomeRegion.fixFixList: sorted by locus (stable), of overlapping
    neighbours the larger was+patch, ties to the first."""
    A = asm()
    assert A.fix_fix_list([(20, "A", "C"), (10, "G", "T")]) == [(10, "G", "T"), (20, "A", "C")]
    # an insertion at a SNP's locus overlaps it: the longer one is kept
    assert A.fix_fix_list([(10, "A", "C"), (10, "", "GGG")]) == [(10, "", "GGG")]
    # equal size: the first in sorted order
    assert A.fix_fix_list([(10, "A", "C"), (10, "", "G")]) == [(10, "A", "C")]
    # a big fix over a small one: kept once, and what follows it survives
    assert A.fix_fix_list([(5, "NNNN", "ACGTACGT"), (7, "N", ""), (12, "A", "T")]) == \
        [(5, "NNNN", "ACGTACGT"), (12, "A", "T")]
    # adjacent, not overlapping
    assert A.fix_fix_list([(5, "AC", ""), (7, "G", "T")]) == [(5, "AC", ""), (7, "G", "T")]


def test_fix_issues_applies_right_to_left():
    A = asm()
    assert A.fix_issues("AAAACCCCGGGG", 101, [(101, "A", "T"), (105, "CCCC", ""), (110, "", "TT")]) == \
        "TAAAGTTGGG"


def test_log_fix_messages():
    """GenomeRegion.logFix's outcome for each kind of result."""
    A = asm()
    assert A.log_fix("c", (10, 19), (0, "", ""), 0, "") == ("NoSolution", "fix break: c:10-19 0 -0 +0 NoSolution\n")
    assert A.log_fix("c", (10, 19), (15, "", ""), 0, "")[0] == "NoChange"
    assert A.log_fix("c", (10, 19), (15, "AC", "GGT"), 0, "")[0] == "BreakFix"
    assert A.log_fix("c", (10, 19), (15, "AC", "GNNT"), 0, "")[0] == "OpenedGap"
    assert A.log_fix("c", (10, 19), (10, "N" * 10, "ACGTACGTAC"), 10, "")[0] == "ClosedGap"
    assert A.log_fix("c", (10, 19), (10, "N" * 10, "ACGTNNNNAC"), 10, "AC") == \
        ("TandemRepeat", "fix gap: c:10-19 10 -0 +6 PartialFill TandemRepeat 2\n")


def test_two_gaps_and_two_deletions():
    """Several attempts in one region, each closed independently."""
    c = AC.planted_region(seed=33, n=20000, gap=(4000, 120), deletion=(15000, 25), snps=(1000, 18000))
    d = c["draft"]
    g2 = 9000
    d = d[:g2] + "N" * 60 + d[g2 + 60:]
    c["draft"] = d
    r = run(c)
    t = c["truth"]
    assert r["status"] == 0 and r["bases"] == t
    assert [l.split()[-1] for l in r["log"].splitlines()] == ["ClosedGap", "ClosedGap", "BreakFix"]
