"""The oracle's tracks (Tracks.standardTracks, Pass 2 of
GenomeRegion.postProcess) on closed forms and brute force."""
import numpy as np
import pytest

from pilon_amd.abi import default_config


def oracle():
    from oracle import oracle as O
    return O


@pytest.mark.parametrize("n", [1, 150, 200, 201, 202, 5000])
def test_smooth_constant(n):
    """Constant c: the window sum of GenomeRegion.smooth holds input(0) beside
    the last 200 values (201c), so the interior is (201c + 100) / 200; the
    first 101 values stay 0 (result(100) is read before it is written); the
    right edge repeats result(n - 101); n <= 200 gives (n c) / n; n = 201
    computes nothing, so every value copies result(100) = 0."""
    c = 37
    out = oracle().smooth(np.full(n, c, np.int32))
    if n <= 200:
        assert (out == (n * c) // n).all()
        return
    if n == 201:                  # n > window, but no index passes i > window: all edges copy result(100) = 0
        assert (out == 0).all()
        return
    interior = (201 * c + 100) // 200
    assert (out[:101] == 0).all()
    assert (out[101:] == interior).all()


def test_smooth_first_value_never_leaves():
    """Only input(0) = 1000: a window that dropped it would give 0."""
    v = np.zeros(3000, np.int32)
    v[0] = 1000
    out = oracle().smooth(v)
    assert (out[101:] == (1000 + 100) // 200).all() and (out[:101] == 0).all()


def test_smooth_random_against_definition():
    rng = np.random.default_rng(5)
    v = rng.integers(0, 300, 4000).astype(np.int32)
    out = oracle().smooth(v)
    cs = np.concatenate([[0], np.cumsum(v.astype(np.int64))])
    for r in range(101, 4000 - 100):
        i = r + 100
        acc = v[0] + cs[i + 1] - cs[i - 199]
        assert out[r] == (acc + 100) // 200
    assert (out[4000 - 100:] == out[4000 - 101]).all()


def brute_gc(contig, start, stop):
    """GC of each locus by the rule of computeGc: a base's value is 1 for G/C,
    0 for A/T, else the value 100 loci earlier (the buffer slot it
    overwrites; alternating 1,0 before the contig); the count at locus L is
    the sum over L-99..L, stored at center L-50 (L itself for the last 50)
    when that number lies in [start, stop]."""
    n = len(contig)
    val = np.zeros(n, np.int64)
    for L in range(n):
        c = contig[L]
        val[L] = 1 if c in "GC" else 0 if c in "AT" else (val[L - 100] if L >= 100 else int(L % 2 == 0))
    init = [int(k % 2 == 0) for k in range(100)]
    out = np.zeros(stop - start + 1, np.int64)
    for L in range(n):
        win = [val[j] if j >= 0 else init[j + 100] for j in range(L - 99, L + 1)]
        cnt = sum(win)
        center = L - 50 if L >= 50 else L
        if start <= center <= stop:
            out[center - start] = cnt
        if start <= L <= stop and L >= n - 50:
            out[L - start] = cnt
    return out


@pytest.mark.parametrize("start,stop", [(1, 700), (120, 400), (650, 700)])
def test_gc_against_brute_force(start, stop):
    rng = np.random.default_rng(start)
    contig = "".join(rng.choice(list("ACGTNacgtR"), p=[.2, .2, .2, .2, .08, .03, .03, .02, .02, .02], size=700))
    got = oracle().gc_track(contig, start, stop)
    assert np.array_equal(got.astype(np.int64), brute_gc(contig, start, stop))


def test_tracks_on_c1_are_consistent():
    """The per-locus tracks against the fields and calls they restate."""
    O = oracle()
    from pilon_amd.synth import SynthData, preset
    from pilon_amd.abi import PL_CF_DISP_MASK, PL_CF_DISP_SHIFT
    d = SynthData(preset("c1"))
    L = len(d.draft(0))
    r = O.run_region(default_config(), d.draft(0), 1, L, [d.batch(0)], name="c", vcf=O.TRACKS)
    t = r["tracks"]
    disp = (r["calls"]["flags"] & PL_CF_DISP_MASK) >> PL_CF_DISP_SHIFT
    assert np.array_equal(t[0], np.isin(disp, [2, 3, 4]).astype(np.int32))
    assert np.array_equal(t[1], (disp != 1).astype(np.int32))
    assert np.array_equal(t[3], r["coverage"]) and np.array_equal(t[4], r["badCoverage"])
    assert np.array_equal(t[8], r["physCoverage"]) and np.array_equal(t[9], r["clips"])
    assert np.array_equal(t[10], r["weightedQual"]) and np.array_equal(t[11], r["weightedMq"])
    good, bad = r["coverage"].astype(np.int64), r["badCoverage"].astype(np.int64)
    assert np.array_equal(t[5], np.where(good + bad > 0, bad * 100 // np.maximum(good + bad, 1), 0))
    assert np.array_equal(t[12], O.gc_track(d.draft(0), 1, L))
    # uniform 40x coverage: copy number 1 (track 0) away from the zeroed left edge
    assert (t[2][:101] == -1).all() and (t[2][200:-200] == 0).mean() > 0.99


def test_fragment_tracks_from_the_final_pileup():
    """Delta, dip and copy number of c1 (which has Pass-1 deletions) from
    fragment coverage rebuilt another way: the final depth, less, at a
    deleted locus, the deletion count of the nearest undeleted locus to its
    left (the one whose call added it).  The oracle keeps processBam's
    depth from before Pass 1 instead."""
    O = oracle()
    from pilon_amd.abi import F
    from pilon_amd.synth import SynthData, preset
    d = SynthData(preset("c1"))
    L = len(d.draft(0))
    r = O.run_region(default_config(), d.draft(0), 1, L, [d.batch(0)], name="c", vcf=O.TRACKS)
    t, pile = r["tracks"], r["pile"].astype(np.int64)
    deleted = (r["calls"]["flags"] & 0x0800) != 0
    assert deleted.sum() > 0
    depth = pile[F["count_A"]] + pile[F["count_C"]] + pile[F["count_G"]] + pile[F["count_T"]] + pile[F["deletions"]]
    frag = depth.copy()
    for i in np.nonzero(deleted)[0]:
        k = i - 1
        while deleted[k]:
            k -= 1
        frag[i] -= pile[F["deletions"]][k]
    assert (frag != depth).any()
    n = L
    idx = np.arange(n)
    left, right = frag[np.maximum(0, idx - 100)], frag[np.minimum(n - 1, idx + 100)]
    assert np.array_equal(t[6], np.abs(left - right))
    assert np.array_equal(t[7], (left - frag) + (right - frag))
    mean = frag.sum() / n
    x = O.smooth(frag.astype(np.int32)) / mean
    cn = (np.floor(x) + ((x - np.floor(x)) >= 0.5)).astype(np.int64)
    assert np.array_equal(t[2], cn - 1)


def jumps_split(d, ci=0):
    """A frags set F and a jumps set J out of one synthetic batch: J takes
    every other read starting in the middle third of the contig, so F's
    depth dips there while the pileup of F+J is the batch's."""
    b = d.batch(ci)
    L = len(d.draft(ci))
    idx = np.arange(b.n)
    jsel = (b.pos > L // 3) & (b.pos < 2 * L // 3) & (idx % 2 == 1)
    return b, b.take(idx[~jsel]), b.take(idx[jsel])


def test_jumps_left_out_of_fragment_coverage():
    """processBam (GenomeRegion.scala:287-300) adds a BAM's depth to fragment
    coverage unless it is a jumps BAM.  Pileups add across BAMs, so a region
    given a frags set F and a jumps set J has the pileup, calls and coverage
    of F+J as one set, and the delta, dip and copy-number tracks (functions
    of fragment coverage alone) of F alone, in either order; an unpaired
    BAM counts as frags."""
    O = oracle()
    from pilon_amd.synth import SynthData, preset
    d = SynthData(preset("c1"))
    draft, L = d.draft(0), len(d.draft(0))
    b, Fs, Js = jumps_split(d)
    assert Js.n > 1000 and Fs.n + Js.n == b.n
    cfg = default_config()
    full = O.run_region(cfg, draft, 1, L, [b], name="c", vcf=O.TRACKS)
    fonly = O.run_region(cfg, draft, 1, L, [Fs], name="c", vcf=O.TRACKS)
    frag_tracks = [2, 6, 7]
    for sets, types in (([Fs, Js], [0, 1]), ([Js, Fs], [1, 0])):
        r = O.run_region(cfg, draft, 1, L, sets, name="c", vcf=O.TRACKS, set_types=types)
        assert np.array_equal(r["pile"], full["pile"])
        assert np.array_equal(r["calls"], full["calls"])
        other = [t for t in range(len(O.TRACK_NAMES)) if t not in frag_tracks]
        assert np.array_equal(r["tracks"][other], full["tracks"][other])
        assert np.array_equal(r["tracks"][frag_tracks], fonly["tracks"][frag_tracks])
        assert not np.array_equal(r["tracks"][6], full["tracks"][6])
    for types in ([0, 0], [0, 2], [2, 2]):
        r = O.run_region(cfg, draft, 1, L, [Fs, Js], name="c", vcf=O.TRACKS, set_types=types)
        assert np.array_equal(r["tracks"], full["tracks"])
    # jumps only: no fragment coverage, so copy number 0 (track value -1)
    # and flat delta and dip
    r = O.run_region(cfg, draft, 1, L, [b], name="c", vcf=O.TRACKS, set_types=[1])
    assert (r["tracks"][2] == -1).all() and not r["tracks"][6].any() and not r["tracks"][7].any()
