"""The per-locus tracks computed on the device (pl_batch_tracks,
pl_region_tracks, pl_batch_wig) against the oracle's Tracks.standardTracks
values, exactly: the known-answer cases (region API), synthetic batches of
many regions, chunk regions of one contig (computeGc's 1-based bounds on
0-based loci, the contig's last 50 loci), regions of at most 201 loci
(smooth's edge cases), a region without reads, and contigs with long
N runs, IUPAC codes and lower case across the GC column-scan chunks."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests import kat_cases as K

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


def same_tracks(got, ref, what):
    from oracle import oracle as O
    assert got.shape == ref.shape, what
    for t in range(ref.shape[0]):
        bad = np.nonzero(got[t] != ref[t])[0]
        assert bad.size == 0, (f"{what}: track {O.TRACK_NAMES[t]} differs at {bad.size} loci, first {bad[0]}: "
                               f"{got[t][bad[0]]} != {ref[t][bad[0]]}")


def wig_text(track, parts):
    """Tracks.makeTrack's text (Tracks.scala:160-177) from the oracle's values."""
    from pilon_amd.engine import TRACK_NAMES, TRACK_FILES
    head = 'track type=wiggle_0 graphType=line color=0,0,255 altColor=255,0,0 name="%s"' % TRACK_NAMES[track]
    if TRACK_FILES[track] == "GC.wig":
        head += " graphType=heatmap midRange=35:65 midColor=0,255,0"
    out = [head + "\n"]
    for name, start, vals in parts:
        out.append("fixedStep chrom=%s start=%d step=1\n" % (name, start))
        out.append("".join("%d\n" % v for v in vals[track].tolist()))
    return "".join(out)


@pytest.mark.parametrize("make", K.ALL, ids=[f.__name__ for f in K.ALL])
def test_tracks_kat(engine, make):
    from oracle import oracle as O
    from pilon_amd.engine import Context, PileUpRegion
    case = make()
    print(f"case: tracks {make.__name__}", flush=True)
    cfg = default_config(**case.get("cfg", {}))
    own = Context(cfg, device=0) if case.get("cfg") else None
    ctx = own or engine
    try:
        ctx.ref_upload(160, "ctg", case["ref"].encode())
        pr = PileUpRegion(ctx, 160, case["start"], case["stop"])
        pr.addReads(K.batch_of(case))
        pr.postProcess()
        got, _ = pr.tracks()
        wig, _ = pr.wig(3)
        # after another batch on the context the region's batch runs again
        ctx.ref_upload(161, "other", case["ref"].encode())
        pr2 = PileUpRegion(ctx, 161, case["start"], case["stop"])
        pr2.addReads(K.batch_of(case))
        pr2.postProcess()
        again, _ = pr.tracks()
        pr2.close()
        pr.close()
    finally:
        if own is not None:
            own.close()
    ref = O.run_region(cfg, case["ref"], case["start"], case["stop"], [K.batch_of(case)], name="ctg", vcf=O.TRACKS)
    same_tracks(got, ref["tracks"], make.__name__)
    assert np.array_equal(again, got)
    assert wig == wig_text(3, [("ctg", case["start"], ref["tracks"])])


def with_gaps(draft, rng):
    """The draft with N runs (one longer than a column-scan chunk of 102400
    loci in contigs of 200 kb or more), IUPAC codes and lower-case stretches."""
    b = bytearray(draft)
    L = len(b)
    if L >= 200000:
        s = L // 3
        b[s:s + 120000] = b"N" * 120000
    for _ in range(max(1, L // 20000)):
        s = int(rng.integers(0, max(1, L - 500)))
        n = int(rng.integers(1, 400))
        b[s:s + n] = bytes(rng.choice(list(b"NRYKMSWnacgt"), n).tolist())
    return bytes(b)


@pytest.mark.parametrize("name,scale,extra,gaps", [
    ("c1", 1.0, {}, False),
    ("c1", 1.0, {}, True),
    ("c2", 0.04, {"n_rate": 0.001, "err_indel_rate": 0.002}, True),
    ("c3", 0.003, {}, False),
    ("c5", 0.01, {}, False),
])
def test_tracks_synthetic_batch(engine, name, scale, extra, gaps):
    from oracle import oracle as O
    from pilon_amd.engine import NTRACKS
    from pilon_amd.synth import SynthData, preset
    print(f"case: tracks {name} x{scale} gaps={gaps}", flush=True)
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    rng = np.random.default_rng(11)
    drafts = [with_gaps(d.draft(ci), rng) if gaps else d.draft(ci) for ci in range(d.n)]
    regions = []
    for ci in range(d.n):
        engine.ref_upload(600 + ci, d.names[ci], drafts[ci])
        rid = engine.readset_upload(600 + ci, d.batch(ci))
        L = len(drafts[ci])
        if name == "c5":   # chunk regions of a contig, as GenomeFile.contigRegions cuts them
            c = (L + 2) // 3
            regions += [(600 + ci, 1 + k * c, min(L, (k + 1) * c), rid) for k in range(3)]
        else:
            regions.append((600 + ci, 1, L, rid))
    if name == "c1":       # smooth's edges: n <= 200, n = 201, 202; a region ending at the contig's end
        L = len(drafts[0])
        regions += [(600, 1, 150, regions[0][3]), (600, 1000, 1200, regions[0][3]),
                    (600, 3000, 3199, regions[0][3]), (600, L - 201, L, regions[0][3])]
    engine.batch_run(regions, window=0, block=0, persistent=1, copy_out=False, want_stats=False)
    got, ms = engine.batch_tracks()
    parts, refs, deleted = [], [], 0
    for cid, s, e, _ in regions:
        ci = cid - 600
        r = O.run_region(engine.config, drafts[ci], s, e, [d.batch(ci)], name=d.names[ci], vcf=O.TRACKS)
        assert r["status"] == 0
        refs.append(r["tracks"])
        parts.append((d.names[ci], s, r["tracks"]))
        deleted += int(np.count_nonzero(r["calls"]["flags"] & 0x0800))
    ref = np.concatenate(refs, axis=1)
    assert got.shape == (NTRACKS, ref.shape[1])
    same_tracks(got, ref, name)
    assert ms > 0
    if name == "c1":
        assert deleted > 0        # Pass 1 deletions: fragment coverage differs from the final depth there
    for t in (2, 12):             # copy number and GC as wig text
        wig, _ = engine.batch_wig(t)
        assert wig == wig_text(t, parts)


def test_tracks_region_without_reads(engine):
    from oracle import oracle as O
    from pilon_amd.reads import ReadBatch
    from pilon_amd.synth import SynthData, preset
    print("case: tracks empty region", flush=True)
    d = SynthData(preset("c1", scale=0.1))
    draft = d.draft(0)
    empty = ReadBatch(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint16), np.zeros(0, np.uint8),
                      np.zeros(1, np.int64), np.zeros(0, np.uint32), np.zeros(1, np.int64), np.zeros(8, np.uint8),
                      np.zeros(8, np.int8))
    engine.ref_upload(700, "empty", draft)
    rid = engine.readset_upload(700, empty)
    engine.batch_run([(700, 1, 3000, rid), (700, 2001, 2100, rid)], window=0, block=0, persistent=1,
                     copy_out=False, want_stats=False)
    got, _ = engine.batch_tracks()
    ref = np.concatenate([O.run_region(engine.config, draft, 1, 3000, [empty], name="empty", vcf=O.TRACKS)["tracks"],
                          O.run_region(engine.config, draft, 2001, 2100, [empty], name="empty",
                                       vcf=O.TRACKS)["tracks"]], axis=1)
    same_tracks(got, ref, "empty")
    assert (got[1] == 1).all() and (got[2] == -1).all()
