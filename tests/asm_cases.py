"""Synthetic reassembly cases: a random truth, a draft made from it with a
gap (a run of N) or a deletion, and error-free reads tiled over the truth.
Each case is (name, reads, draft, region start, region stop, break start,
break stop, is_gap, options, truth, extra facts), loci 1-based as in Pilon.
"""
import numpy as np

from pilon_amd.reads import pack_reads


def random_bases(rng, n):
    return "".join("ACGT"[c] for c in rng.integers(0, 4, n))


def tiled_reads(truth, step=5, length=150, qual=30, mapq=60, holes=(), alleles=None):
    """Reads of `truth` every `step` bases, `length` long; a read that
    overlaps a hole (0-based [a, b)) is left out; alleles {position: base}
    go into every other read that covers the position."""
    reads, k = [], 0
    for s in range(0, len(truth) - length + 1, step):
        if any(s < b and s + length > a for a, b in holes):
            continue
        seq = truth[s:s + length]
        if alleles and k % 2:
            seq = list(seq)
            for p, b in alleles.items():
                if s <= p < s + length:
                    seq[p - s] = b
            seq = "".join(seq)
        k += 1
        reads.append(dict(pos=s + 1, cigar="%dM" % length, seq=seq, qual=[qual] * length, flag=0, mapq=mapq))
    return pack_reads(reads)


def gap_case(seed=1, n=2000, gap=(801, 900), step=5, holes=(), snps=(), n_len=None, **opts):
    """Draft = truth with bases gap[0]..gap[1] (1-based) replaced by a run
    of n_len N (the gap's own length by default); every other read carries
    the next base (ACGT order) at each 0-based position in snps."""
    rng = np.random.default_rng(seed)
    truth = random_bases(rng, n)
    a, b = gap
    n_len = b + 1 - a if n_len is None else n_len
    draft = truth[:a - 1] + "N" * n_len + truth[b:]
    alleles = {p: "ACGT"[("ACGT".index(truth[p]) + 1) % 4] for p in snps}
    reads = tiled_reads(truth, step=step, holes=holes, alleles=alleles)
    return dict(name="gap%d" % seed, reads=reads, draft=draft, start=1, stop=len(draft), bstart=a,
                bstop=a + n_len - 1, is_gap=True, opts=opts, truth=truth)


def deletion_case(seed=2, n=2000, at=1000, missing=40, brk=(991, 1010), **opts):
    """Draft = truth without truth[at:at+missing] (0-based); the break
    [brk] straddles the junction.  The base after the deletion differs from
    the first deleted one, so the insertion has a single left-most place."""
    rng = np.random.default_rng(seed)
    while True:
        truth = random_bases(rng, n)
        if missing == 0 or truth[at] != truth[at + missing]:
            break
    draft = truth[:at] + truth[at + missing:]
    reads = tiled_reads(truth)
    return dict(name="del%d" % seed, reads=reads, draft=draft, start=1, stop=len(draft), bstart=brk[0],
                bstop=brk[1], is_gap=False, opts=opts, truth=truth, at=at, missing=missing)


def tandem_case(seed=3, unit=20, copies=6, n=2000, gap=(801, 900), **opts):
    """A tandem repeat of `copies` units inside the gap: longer than K plus
    one unit, so the K-mer walk comes back to a K-mer it holds."""
    rng = np.random.default_rng(seed)
    u = random_bases(rng, unit)
    truth = random_bases(rng, 820) + u * copies + random_bases(rng, n - 820 - unit * copies)
    a, b = gap
    draft = truth[:a - 1] + "N" * (b + 1 - a) + truth[b:]
    reads = tiled_reads(truth)
    return dict(name="tr%d" % seed, reads=reads, draft=draft, start=1, stop=n, bstart=a, bstop=b, is_gap=True,
                opts=opts, truth=truth, unit=u)


def fixed_base_case(which):
    """gap_case(1) with the base after the gap changed: in the region's
    current bases only ("bases", a fix made earlier in the region) or in
    the contig as read too ("both")."""
    c = gap_case(1)
    d = c["draft"]
    x = "ACGT"[("ACGT".index(d[900]) + 2) % 4]
    c["draft"] = d[:900] + x + d[901:]
    if which == "bases":
        c["contig"] = d
    c["name"] = "fixed_" + which
    c["x"] = x
    return c


def parity_cases():
    """The cases the device path is checked on against the oracle."""
    out = [gap_case(1), gap_case(4, holes=((830, 860),)), gap_case(16, holes=((870, 880),)),
           gap_case(17, gap=(801, 950), n_len=100), gap_case(18, gap=(801, 950), n_len=100, gap_margin=20),
           gap_case(5, snps=(850, 870)),
           gap_case(6, snps=(850,), multi_closure=1), deletion_case(2), deletion_case(7, missing=3),
           deletion_case(8, fix_breaks=1), tandem_case(3), tandem_case(9, tr_safe=0), tandem_case(9, unit=10, copies=7, tr_safe=0, multi_closure=1), tandem_case(10, unit=7, copies=12),
           gap_case(11, step=40), gap_case(12, gap=(1001, 1400), break_radius=300),
           gap_case(13, gap=(801, 1000), gap_margin=50)]
    c = gap_case(14)
    c["reads"] = pack_reads([])
    c["name"] = "no_reads"
    out.append(c)
    c = deletion_case(15, missing=0)
    c["name"] = "no_change"
    out.append(c)
    return out


def planted_region(seed=31, n=12000, gap=(3000, 100), deletion=(8000, 40), snps=(1500, 5500, 10500), step=5,
                   length=150, unmapped_in_gap=False, truth=None):
    """A draft made from a random truth with SNPs (the next base, ACGT
    order), a run of N standing for truth[gap[0]:gap[0]+gap[1]] (0-based,
    same length) and truth[deletion[0]:deletion[0]+deletion[1]] left out,
    with error-free reads of the truth every `step` bases, alternately
    reverse, aligned to the draft as an aligner would: M across the N run,
    and across the deletion the longer side as M with the rest soft-clipped.
    Returns dict(truth, draft, reads, loci of the draft's features); a
    truth given is used as it is."""
    rng = np.random.default_rng(seed)
    d0, dl = deletion
    while truth is None:
        truth = random_bases(rng, n)
        if truth[d0] == truth[d0 + dl]:
            truth = None
    n = len(truth)
    g0, gl = gap
    t = list(truth)
    for p in snps:
        t[p] = "ACGT"[("ACGT".index(truth[p]) + 1) % 4]
    t[g0:g0 + gl] = "N" * gl
    draft = "".join(t[:d0] + t[d0 + dl:])
    reads = []
    for k, s in enumerate(range(0, n - length + 1, step)):
        seq = truth[s:s + length]
        flag = 0x10 if k % 2 else 0
        e = s + length
        if e <= d0:
            pos, cig = s + 1, "%dM" % length
        elif s >= d0 + dl:
            pos, cig = s - dl + 1, "%dM" % length
        else:
            left, right = max(d0 - s, 0), max(e - (d0 + dl), 0)
            if left >= right:
                pos, cig = s + 1, "%dM%dS" % (left, length - left)
            else:
                pos, cig = d0 + 1, "%dS%dM" % (length - right, right)
        if unmapped_in_gap and s >= g0 and e <= g0 + gl:
            continue
        reads.append(dict(pos=pos, cigar=cig, seq=seq, qual=[30] * length, flag=flag, mapq=60))
    return dict(truth=truth, draft=draft, reads=pack_reads(reads), gap=(g0 + 1, g0 + gl),
                junction=d0 + 1, snps=[p + 1 for p in snps])
