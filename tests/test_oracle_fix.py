"""GenomeFile's region list and output loop: the oracle's restatements of
contigRegions, parseTargets, writeFastaElement and processRegions' newName
pinned by hand-derived answers, then the library's host entry points
(pl_contig_regions, pl_parse_targets, pl_fasta_element, pl_pilon_name; no
GPU needed) against them on random inputs."""
import os
import random
import tempfile

import pytest

from pilon_amd import engine as E


def oracle():
    from oracle import oracle as O
    return O


# ------------------------------------------------------------------ pins
def test_contig_regions_known():
    """GenomeFile.contigRegions with chunkSize 10,000,000: a 25 Mb contig is
    3 chunks of ceil(25e6 / 3) = 8,333,334 (the last one shorter); a contig
    of exactly one chunk is one region; 20,000,001 bases need 3 chunks of
    6,666,667."""
    O = oracle()
    assert O.contig_regions(25_000_000) == [(1, 8333334), (8333335, 16666668), (16666669, 25000000)]
    assert O.contig_regions(10_000_000) == [(1, 10_000_000)]
    assert O.contig_regions(20_000_001) == [(1, 6666667), (6666668, 13333334), (13333335, 20000001)]
    assert O.contig_regions(1) == [(1, 1)]
    assert O.contig_regions(10, 3) == [(1, 3), (4, 6), (7, 9), (10, 10)]   # nChunks 4, chunk ceil(10/4) = 3


def test_contig_regions_invariants():
    """Regions tile [1, len] in order, at most nChunks = ceil(len / chunk) of
    them, none longer than chunk."""
    O = oracle()
    rng = random.Random(3)
    for _ in range(2000):
        L = rng.randint(1, 10 ** 6)
        ch = rng.randint(1, 10 ** 5)
        r = O.contig_regions(L, ch)
        assert r[0][0] == 1 and r[-1][1] == L
        assert all(b[0] == a[1] + 1 for a, b in zip(r, r[1:]))
        assert len(r) == -(-L // ch) and max(e - s + 1 for s, e in r) <= ch


def test_pilon_name_known():
    """GenomeFile.scala:137-140: "_pilon" for a plain name, "|pilon" when the
    name holds a '|', "pilon" when it ends with one."""
    O = oracle()
    assert O.pilon_name("contig_7") == "contig_7_pilon"
    assert O.pilon_name("gi|5524211|gb|AAD44166.1") == "gi|5524211|gb|AAD44166.1|pilon"
    assert O.pilon_name("gi|5524211|") == "gi|5524211|pilon"
    assert O.pilon_name("|") == "|pilon"


def test_fasta_element_known():
    """writeFastaElement: lines of 80, the last shorter, none for nothing."""
    O = oracle()
    assert O.fasta_element("a", "ACGT") == ">a\nACGT\n"
    assert O.fasta_element("a", "A" * 80) == ">a\n" + "A" * 80 + "\n"
    assert O.fasta_element("a", "C" * 160) == ">a\n" + "C" * 80 + "\n" + "C" * 80 + "\n"
    assert O.fasta_element("a", "G" * 161) == ">a\n" + "G" * 80 + "\n" + "G" * 80 + "\nG\n"
    assert O.fasta_element("a", "") == ">a\n"


CONTIGS = [("chr1", 1000), ("chr2", 500), ("chr1:1-3", 700), ("gi|7|", 90)]


def test_parse_targets_known():
    """parseTargetString: greedy name before the last ":start-stop" with
    optional surrounding whitespace; no match = a whole contig; comma
    separated, in order, duplicates kept."""
    O = oracle()
    assert O.parse_target_string("chr1:5-10", CONTIGS) == [(0, 5, 10)]
    assert O.parse_target_string("  chr1:5-10 \t", CONTIGS) == [(0, 5, 10)]
    assert O.parse_target_string("chr2", CONTIGS) == [(1, 1, 500)]
    assert O.parse_target_string("chr1:1-3:4-9", CONTIGS) == [(2, 4, 9)]          # name "chr1:1-3"
    assert O.parse_target_string("chr1:1-3", CONTIGS) == [(0, 1, 3)]
    assert O.parse_target_string("chr1:7-7,chr2,chr1:7-7", CONTIGS) == [(0, 7, 7), (1, 1, 500), (0, 7, 7)]
    assert O.parse_target_string("gi|7|:1-90,", CONTIGS) == [(3, 1, 90)]             # trailing empty dropped
    for bad in ["chr9", " chr2", "chr1:5-1001", "chr1:0-5", "chr1:9-5", "chr1:1-99999999999", "chr1:5-10,,chr2"]:
        with pytest.raises(ValueError):
            O.parse_target_string(bad, CONTIGS)


# ------------------------------------------------------- product vs oracle
def test_product_contig_regions_match_oracle():
    O = oracle()
    rng = random.Random(11)
    cases = [(25_000_000, 10_000_000), (10_000_000, 10_000_000), (1, 10_000_000), (2_000_000_000, 10_000_000)]
    cases += [(rng.randint(1, 10 ** 7), rng.randint(1, 10 ** 6)) for _ in range(300)]
    for L, ch in cases:
        assert E.contig_regions(L, ch) == O.contig_regions(L, ch), (L, ch)


def test_product_names_and_fasta_match_oracle():
    O = oracle()
    rng = random.Random(5)
    for _ in range(300):
        name = "".join(rng.choice("ab|_:.1") for _ in range(rng.randint(1, 12)))
        assert E.pilon_name(name) == O.pilon_name(name), name
        seq = "".join(rng.choice("ACGTN") for _ in range(rng.randint(0, 400)))
        assert E.fasta_element(name, seq) == O.fasta_element(name, seq)


def test_product_parse_targets_match_oracle():
    O = oracle()
    rng = random.Random(9)
    pieces = ["chr1", "chr2", "chr1:1-3", "gi|7|", "chr9", "", " ", ":", "-", "5", "10", "1000", "0", ",", "\t"]
    for _ in range(3000):
        s = "".join(rng.choice(pieces) for _ in range(rng.randint(1, 7)))
        try:
            want = O.parse_target_string(s, CONTIGS)
        except ValueError:
            want = None
        try:
            got = E.parse_targets(s, CONTIGS)
        except E.PilonError as e:
            assert e.status == 1
            got = None
        assert got == want, repr(s)


def test_product_parse_targets_file():
    """A readable file whose lines all parse: its lines' targets; otherwise
    the string itself is parsed (GenomeFile.scala:238-250)."""
    O = oracle()
    with tempfile.TemporaryDirectory() as d:
        good = os.path.join(d, "t.txt")
        with open(good, "w") as f:
            f.write("chr1:1-5\nchr2:3-4,chr1\r\ngi|7|\n")
        assert E.parse_targets(good, CONTIGS) == O.parse_targets(good, CONTIGS) == \
            [(0, 1, 5), (1, 3, 4), (0, 1, 1000), (3, 1, 90)]
        bad = os.path.join(d, "chr2")                       # a file whose line fails: the name as a target
        with open(bad, "w") as f:
            f.write("chr9:1-5\n")
        with pytest.raises(E.PilonError):
            E.parse_targets(bad, CONTIGS)
