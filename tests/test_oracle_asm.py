"""The oracle's local reassembly (GapFiller.fillGap / fixBreak over the
Assembler's K-mer graph) against what the construction of each case fixes:
a gap filled with the truth, a deletion put back at its left-most place, a
partial extension into a gap, a tandem repeat noted as one unit, no fix
without reads; and the string helpers on hand cases."""
import pytest

from pilon_amd.abi import default_config

from . import asm_cases as AC


def asm():
    from oracle import asm as A
    return A


def run(c, **over):
    opts = dict(c["opts"], **over)
    d = c["draft"]
    return asm().assemble_break(default_config(), c["reads"], d[c["start"] - 1:c["stop"]], c.get("contig", d),
                                c["start"], c["stop"], c["bstart"], c["bstop"], c["is_gap"], **opts)


def test_proper_overlap_hand_cases():
    """GapFiller.properOverlap: the first overlap of at least minOverlap
    where the left's suffix meets the right's prefix, or the right's
    window inside a longer left."""
    G = asm().GapFiller
    assert G.proper_overlap("ACGTACGTTT", "GTTTCCA", 3) == "ACGTACGTTTCCA"
    assert G.proper_overlap("ACGTACGTTT", "GTTTCCA", 5) == ""
    assert G.proper_overlap("AAACCC", "GGG", 1) == ""
    # the right begins before the left: the join keeps the right's tail
    assert G.proper_overlap("CCCGGGTT", "AACCCGGG", 3) == "CCCGGG"
    # empty left: the first offset compares nothing and matches
    assert G.proper_overlap("", "ACGTACGT", 2) == "GTACGT"


def test_consensus_hand_cases():
    G = asm().GapFiller
    assert G.consensus_from_left(["ACGTA", "ACGAA", "ACG"]) == "ACG"
    assert G.consensus_from_left(["ACGTA", "ACTTA"]) == "AC"
    assert G.consensus_from_left(["ACGTA"]) == "ACGTA"
    assert G.consensus_from_right(["TTACG", "AACG"]) == "ACG"
    assert G.consensus_from_right(["TTACG", "TTACC"]) == ""
    assert G.consensus_from_right(["GGTTACG", "CG"]) == "CG"


def test_kmer_path_string_and_loop_note():
    """kmerPathString joins a newest-first K-mer list; a walk round a
    self-contained cycle notes the cycle's length and its sequence."""
    A = asm().Assembler
    assert A.kmer_path_string(["GTA", "CGT", "ACG"]) == "ACGTA"
    # K = 3 over the cycle ACGTACGT...: ACG -> CGT -> GTA -> TAC -> ACG
    graph = [("ACG", "T", None), ("CGT", "A", None), ("GTA", "C", None), ("TAC", "G", None)]
    a = A(graph, 3)
    paths = a.paths_forward("ACG")
    assert paths == ["ACGTAC"]
    assert a.loop_length == 4 and a.loop_sequence in "ACGTACGT"


def test_branches_are_explored_in_kgraph_order():
    """Two successors: both branches, kGraph's first; equal lengths keep
    that order (pathsForward's sort is stable)."""
    A = asm().Assembler
    graph = [("AAC", "G", "T"), ("ACG", "A", None), ("ACT", "A", None)]
    assert A(graph, 3).paths_forward("AAC") == ["AACGA", "AACTA"]


def test_gap_filled_with_the_truth():
    c = AC.gap_case(1)
    st, sol, loop = run(c)
    assert st == 0 and loop == ""
    assert sol == (801, "N" * 100, c["truth"][800:900])


def test_trim_takes_current_or_original_base():
    """trimPatch stops where the patch differs from both the region's
    current base and the contig's: a base changed in the current bases only
    is trimmed as before; changed in the contig too, it joins the fix."""
    c = AC.fixed_base_case("bases")
    assert run(c) == (0, (801, "N" * 100, c["truth"][800:900]), "")
    c = AC.fixed_base_case("both")
    assert run(c) == (0, (801, "N" * 100 + c["x"], c["truth"][800:901]), "")


def test_longer_gap_wider_radius():
    c = AC.gap_case(12, gap=(1001, 1400), break_radius=300)
    st, sol, loop = run(c)
    assert st == 0 and sol == (1001, "N" * 400, c["truth"][1000:1400])


def test_gap_shorter_than_the_sequence_it_hides():
    """100 N stand for 150 bases: the fill is the 150, gapMargin allowing."""
    c = AC.gap_case(17, gap=(801, 950), n_len=100)
    st, sol, _ = run(c)
    assert st == 0 and sol == (801, "N" * 100, c["truth"][800:950])


def test_gap_margin_rejects_the_closure():
    """With gapMargin 20 the 50-base difference rejects the closure; the
    fallback joins both consensus paths with minGap N (the paths overlap),
    trimmed against the draft: longer than the gap and holding the N run."""
    c = AC.gap_case(18, gap=(801, 950), n_len=100, gap_margin=20)
    st, (start, ref, patch), _ = run(c)
    assert st == 0 and start == 801 and ref == "N" * 100
    assert "N" * 10 in patch and "N" * 11 not in patch and len(patch) > 150
    assert patch.startswith(c["truth"][800:950]) and patch.endswith(c["truth"][800:950])


def test_deletion_put_back():
    for missing in (40, 3):
        c = AC.deletion_case(2 if missing == 40 else 7, missing=missing)
        st, sol, loop = run(c)
        assert st == 0 and loop == ""
        assert sol == (c["at"] + 1, "", c["truth"][c["at"]:c["at"] + missing])


def test_break_with_nothing_to_change():
    """A break whose flanks join back to the draft itself: trimPatch leaves
    (locus, "", ""), which is not the no-solution triple."""
    c = AC.deletion_case(15, missing=0)
    st, (start, ref, patch), _ = run(c)
    assert st == 0 and start > 0 and ref == "" and patch == ""


def test_partial_extension_into_a_gap():
    """Reads stop 70 bases into the gap (none overlap truth 870-879): the
    left path reaches into the gap, the right one does not; the N that the
    fallback puts between them meet the draft's N and are trimmed, so the
    fix replaces the gap's first bases with the truth and keeps the rest."""
    c = AC.gap_case(16, holes=((870, 880),))
    st, (start, ref, patch), loop = run(c)
    assert st == 0 and loop == ""
    assert start == 801 and len(patch) >= 20 and ref == "N" * len(patch)
    assert patch == c["truth"][800:800 + len(patch)]


def test_no_extension_no_fix():
    """Reads missing over 30 bases cut the graph before either walk gets
    minExtend past its anchor: no fix."""
    c = AC.gap_case(4, holes=((830, 860),))
    assert run(c)[:2] == (0, (0, "", ""))


def test_no_reads_no_fix():
    c = AC.gap_case(14)
    c["reads"] = AC.pack_reads([])
    assert run(c) == (0, (0, "", ""), "")


@pytest.mark.parametrize("unit,copies,seed", [(20, 6, 3), (7, 12, 10)])
def test_tandem_repeat_noted(unit, copies, seed):
    """A repeat longer than K plus one unit: the walk meets a K-mer it holds
    one unit back; the loop noted is the unit (some rotation of it).  With
    trSafe the closed path is refused and the gap fallback opens N in it."""
    c = AC.tandem_case(seed, unit=unit, copies=copies)
    st, (start, ref, patch), loop = run(c)
    assert st == 0 and len(loop) == unit and loop in c["unit"] * 2
    assert start == 801 and "N" in patch


@pytest.mark.parametrize("seed,unit,copies", [(9, 10, 7), (21, 12, 6)])
def test_tandem_repeat_collapsed_without_tr_safe(seed, unit, copies):
    """trSafe off, multiClosure on: the K-mer at repeat offset
    (copies*unit - K) mod unit already has the exit as a second successor,
    so the shortest walk leaves the repeat floor((copies*unit - K) / unit)
    units early on each side; the smallest change closes the gap with the
    repeat collapsed by that many units."""
    c = AC.tandem_case(seed, unit=unit, copies=copies)
    st, (start, ref, patch), loop = run(c, tr_safe=0, multi_closure=1)
    t, u = c["truth"], c["unit"]
    lost = (copies * unit - 47) // unit
    assert st == 0 and len(loop) == unit and start == 801 and ref == "N" * 100
    assert t[820:820 + unit * copies] == u * copies
    assert patch == t[800:820] + u * (copies - lost) + t[820 + unit * copies:900]


def test_heterozygous_sites_pick_one_haplotype():
    """Half the reads carry another base at two sites: every combination
    closes with the same length, so one fix (the first found) is kept, and
    it is the truth with each site taking one of its two bases."""
    c = AC.gap_case(5, snps=(850, 870))
    st, (start, ref, patch), _ = run(c)
    t = c["truth"]
    assert st == 0 and start == 801 and ref == "N" * 100 and len(patch) == 100
    for i, (p, b) in enumerate(zip(patch, t[800:900])):
        if 800 + i in (850, 870):
            assert p in (b, "ACGT"[("ACGT".index(b) + 1) % 4])
        else:
            assert p == b


def test_gap_at_the_region_start_is_an_edge():
    """A gap that starts the region: the left flank is empty and Pilon's
    subString require fails."""
    c = AC.gap_case(1)
    c["bstart"] = 1
    st, sol, _ = run(c)
    assert st == 4 and sol == (0, "", "")
