"""The oracle's VCF records (Vcf.writeRecord, Vcf.scala:74-163, as
GenomeRegion.writeVcf calls it) on hand-derived known answers, and Java's
"%.2f" half-up rounding of AF against Python's decimal arithmetic."""
import random
from decimal import ROUND_HALF_UP, Decimal

import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests import kat_cases as K


def oracle():
    from oracle import oracle as O
    return O


def vcf_lines(case, flags=None):
    O = oracle()
    r = O.run_region(default_config(**case.get("cfg", {})), case["ref"], case["start"], case["stop"],
                     [K.batch_of(case)], name="ctg", vcf=O.VCF_ON if flags is None else flags)
    return r["vcf"].splitlines()


def java_af(alt, tot):
    """(altDP.toFloat / tot.toFloat) formatted with "%.2f": the float's exact
    value rounded half up (decimal module, independent of the oracle)."""
    f = np.float32(np.float32(alt) / np.float32(tot))
    return str(Decimal(float(f)).quantize(Decimal("0.01"), rounding=ROUND_HALF_UP))


@pytest.mark.parametrize("af,txt", [(0.0, "0.00"), (1.0, "1.00"), (0.125, "0.13"), (0.375, "0.38"), (0.625, "0.63"),
                                    (0.875, "0.88"), (0.005, "0.00"), (0.015, "0.01"), (0.995, "1.00"), (0.5, "0.50")])
def test_af_half_up_known(af, txt):
    # 0.125 is a tie in binary and decimal: half up, not half even; 0.005 and
    # 0.015 as floats lie just below their ties; 0.995 just above
    assert oracle().fmt_af(af) == txt


def test_af_against_decimal():
    rng = random.Random(7)
    for _ in range(20000):
        tot = rng.randint(1, 2000)
        alt = rng.randint(0, tot)
        f = float(np.float32(np.float32(alt) / np.float32(tot)))
        assert oracle().fmt_af(f) == java_af(alt, tot), (alt, tot)


def test_deletion_records():
    """14M2D16M x6 (case_deletion_call): the DEL call at 15 prints at 14 as
    CAG -> C (REF = refBase(14) + "AG", QUAL "." for an indel deletion),
    delPct 100 -> AF 1.00, 1/1; the same index then prints its second,
    embedded record (Del filter, DP = count = 0); locus 16 is deleted by
    Pass 1 and prints embedded with DC = 6 from the propagation."""
    L = vcf_lines(K.case_deletion_call())
    assert L[14] == "ctg\t14\t.\tCAG\tC\t.\tPASS\tDP=6;TD=6;BQ=0;MQ=60;QD=0;BC=0,0,0,0;QP=0,0,0,0;PC=6;IC=0;DC=6;" \
                    "XC=0;AC=2;AF=1.00\tGT\t1/1"
    assert L[15].startswith("ctg\t15\t.\tA\t.\t0\tDel\tDP=0;TD=6;") and L[15].endswith("AC=0;AF=0.00\tGT\t0/0")
    assert L[16].startswith("ctg\t16\t.\tG\t.\t0\tDel\tDP=0;") and ";DC=6;" in L[16]
    assert L[13] == "ctg\t14\t.\tC\t.\t180\tPASS\tDP=6;TD=6;BQ=30;MQ=60;QD=30;BC=0,6,0,0;QP=0,100,0,0;PC=6;IC=0;" \
                    "DC=0;XC=0;AC=0;AF=0.00\tGT\t0/0"


def test_deletion_after_n_anchor():
    """case_deletion_n_anchor: the deletion record's anchor is the draft's N,
    so REF is "NAG", cBase "N" prints as "." and AF is 0.00 (Vcf.scala:131,
    138-141); QUAL stays "." for a deletion.  Locus 14 itself: homozygous C
    over rBase N prints 1/1 with ALT C."""
    L = vcf_lines(K.case_deletion_n_anchor())
    assert L[14] == "ctg\t14\t.\tNAG\t.\t.\tPASS\tDP=6;TD=6;BQ=0;MQ=60;QD=0;BC=0,0,0,0;QP=0,0,0,0;PC=6;IC=0;DC=6;" \
                    "XC=0;AC=2;AF=0.00\tGT\t1/1"
    assert L[13].startswith("ctg\t14\t.\tN\tC\t180\tPASS\tDP=6;") and L[13].endswith("AC=2;AF=1.00\tGT\t1/1")


def test_ambiguous_record_and_flags():
    """case_ambiguous, locus 20: C and G 5 reads each at q30 MAPQ 60.
    qualSum C = G = 9150, base C (stable order), alt G, rBase T: the het
    branch prints REF T, ALT C, refDP = altBaseSum, altDP = baseSum -> AF 0.50;
    score 300; BQ = roundDiv(18300, roundDiv(610*10, 10)) = 30; MQ =
    roundDiv(610 - 10, 10) = 60; QP = (100*9150 + 9150)/18300 = 50."""
    line = "ctg\t20\t.\tT\tC\t300\t{f}\tDP=10;TD=10;BQ=30;MQ=60;QD=30;BC=0,5,5,0;{q};PC=10;IC=0;DC=0;XC=0;" \
           "AC=1;AF=0.50\tGT\t0/1"
    O = oracle()
    assert vcf_lines(K.case_ambiguous())[19] == line.format(f="Amb", q="QP=0,50,50,0")
    assert vcf_lines(K.case_ambiguous(), O.VCF_ON | O.VCF_DIPLOID)[19] == line.format(f="PASS", q="QP=0,50,50,0")
    assert vcf_lines(K.case_ambiguous(), O.VCF_ON | O.VCF_QE)[19] == line.format(f="Amb", q="QE=0,9150,9150,0")
    # --iupac changes the fix, not the record
    assert vcf_lines(K.case_ambiguous_iupac())[19] == line.format(f="Amb", q="QP=0,50,50,0")


def test_insertion_records():
    """case_insertion_rotation: the insertion called at 13 prints at 12 as
    G -> GAC (QUAL = score, not "."), then its second record at 13."""
    L = vcf_lines(K.case_insertion_rotation())
    i = [k for k, l in enumerate(L) if "\tGAC\t" in l]
    assert len(i) == 1
    ins, nxt = L[i[0]], L[i[0] + 1]
    assert ins.startswith("ctg\t12\t.\tG\tGAC\t192\tPASS\t") and ins.endswith(";IC=6;DC=0;XC=0;AC=2;AF=1.00\tGT\t1/1")
    assert nxt.startswith("ctg\t13\t.\tA\t.\t192\tPASS\t") and nxt.endswith("\tGT\t0/0")


def test_one_record_per_locus_plus_indels():
    for make in K.ALL:
        c = make()
        L = vcf_lines(c)
        assert len(L) >= c["stop"] - c["start"] + 1
        assert all(l.count("\t") == 9 for l in L)


def test_duplication_record():
    """GenomeRegion.writeVcf (GenomeRegion.scala:623-631), Vcf.writeDup
    (Vcf.scala:194-202): the two-BAM case whose second BAM doubles 15 kb has
    one duplicationEvents region -- the bed's Duplication line, whose region
    derivation is pinned in test_oracle_bed -- announced by one line before
    the records of its first locus: refBase(start - 1), SVLEN = size, END =
    stop, GT ./. .  Every other line is a locus record."""
    from tests.test_oracle_bed import dup_case
    O = oracle()
    d, draft, sets = dup_case()
    r = O.run_region(default_config(), draft, 1, len(draft), sets, name=d.names[0], vcf=O.VCF_ON | O.BED)
    dups = [l.split("\t") for l in r["bed"].splitlines() if l.split("\t")[3] == "Duplication"]
    assert len(dups) == 1
    a, b = int(dups[0][1]) + 1, int(dups[0][2])        # 1-based start, stop of the event
    assert b - a + 1 > 10000
    L = r["vcf"].splitlines()
    k = [i for i, l in enumerate(L) if "<DUP>" in l]
    assert len(k) == 1
    want = f"{d.names[0]}\t{a - 1}\t.\t{chr(draft[a - 2]).upper()}\t<DUP>\t.\tPASS\t" \
           f"SVTYPE=DUP;SVLEN={b - a + 1};END={b};IMPRECISE\tGT\t./."
    assert L[k[0]] == want
    assert L[k[0] + 1].split("\t")[1] == str(a)        # then the records of locus a
    assert L[k[0] - 1].split("\t")[1] == str(a - 1)
    assert len(L) >= len(draft) + 1
    # without the doubled reads there is none
    r2 = O.run_region(default_config(), draft, 1, len(draft), sets[:1], name=d.names[0], vcf=O.VCF_ON)
    assert "<DUP>" not in r2["vcf"]
