"""Pilon.bed from the device (pl_batch_bed, pl_region_bed) against the
oracle's lines, exactly: the known-answer cases, synthetic batches (N runs,
chunk regions of one contig, regions of a few hundred loci), a second BAM
that doubles 15 kb (copy number 2: Copy# and Duplication regions) beside an
N gap, and regions without reads."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests import kat_cases as K

pytestmark = pytest.mark.gpu

HEAD = 'track description="Issues found by Pilon" name="Pilon"\n'


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


def first_diff(a, b):
    la, lb = a.splitlines(), b.splitlines()
    for k, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return k, x, y
    return min(len(la), len(lb)), len(la), len(lb)


@pytest.mark.parametrize("make", K.ALL, ids=[f.__name__ for f in K.ALL])
def test_bed_kat(engine, make):
    from oracle import oracle as O
    from pilon_amd.engine import Context, PileUpRegion
    case = make()
    print(f"case: bed {make.__name__}", flush=True)
    cfg = default_config(**case.get("cfg", {}))
    own = Context(cfg, device=0) if case.get("cfg") else None
    ctx = own or engine
    try:
        ctx.ref_upload(170, "ctg", case["ref"].encode())
        pr = PileUpRegion(ctx, 170, case["start"], case["stop"])
        pr.addReads(K.batch_of(case))
        pr.postProcess()
        got, _ = pr.bed()
        pr.close()
    finally:
        if own is not None:
            own.close()
    ref = O.run_region(cfg, case["ref"], case["start"], case["stop"], [K.batch_of(case)], name="ctg", vcf=O.BED)
    assert got == HEAD + ref["bed"]


@pytest.mark.parametrize("name,scale,extra,gaps", [
    ("c1", 1.0, {}, False),
    ("c2", 0.04, {"n_rate": 0.001, "err_indel_rate": 0.002}, True),
    ("c3", 0.003, {}, False),
    ("c5", 0.01, {}, False),
])
def test_bed_synthetic_batch(engine, name, scale, extra, gaps):
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    from tests.test_gpu_tracks import with_gaps
    print(f"case: bed {name} x{scale} gaps={gaps}", flush=True)
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    rng = np.random.default_rng(13)
    drafts = [with_gaps(d.draft(ci), rng) if gaps else d.draft(ci) for ci in range(d.n)]
    regions = []
    for ci in range(d.n):
        engine.ref_upload(800 + ci, d.names[ci], drafts[ci])
        rid = engine.readset_upload(800 + ci, d.batch(ci))
        L = len(drafts[ci])
        if name == "c5":
            c = (L + 2) // 3
            regions += [(800 + ci, 1 + k * c, min(L, (k + 1) * c), rid) for k in range(3)]
        else:
            regions.append((800 + ci, 1, L, rid))
    if name == "c1":
        L = len(drafts[0])
        regions += [(800, 1, 150, regions[0][3]), (800, 1000, 1450, regions[0][3]), (800, L - 700, L, regions[0][3])]
    engine.batch_run(regions, window=0, block=0, persistent=1, copy_out=False, want_stats=False)
    got, ms = engine.batch_bed(track_line=False)
    ref = "".join(O.run_region(engine.config, drafts[cid - 800], s, e, [d.batch(cid - 800)], name=d.names[cid - 800],
                               vcf=O.BED)["bed"] for cid, s, e, _ in regions)
    assert got == ref, first_diff(got, ref)
    assert ref and ms > 0


def test_bed_duplication_and_gap(engine):
    from oracle import oracle as O
    from pilon_amd.engine import PileUpRegion
    from tests.test_oracle_bed import dup_case
    print("case: bed duplication + gap, two BAMs", flush=True)
    d, draft, sets = dup_case()
    engine.ref_upload(900, d.names[0], draft)
    pr = PileUpRegion(engine, 900, 1, len(draft))
    for b in sets:
        pr.addReads(b)
    pr.postProcess()
    got, _ = pr.bed(track_line=False)
    pr.close()
    ref = O.run_region(engine.config, draft, 1, len(draft), sets, name=d.names[0], vcf=O.BED)["bed"]
    assert got == ref, first_diff(got, ref)
    assert {"Copy#", "Duplication", "Gap"} <= {l.split("\t")[3] for l in got.splitlines()}


def test_bed_regions_without_reads(engine):
    from oracle import oracle as O
    from pilon_amd.reads import ReadBatch
    from pilon_amd.synth import SynthData, preset
    print("case: bed empty regions", flush=True)
    d = SynthData(preset("c1", scale=0.1))
    draft = bytearray(d.draft(0))
    draft[1500:1530] = b"N" * 30
    draft = bytes(draft)
    empty = ReadBatch(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint16), np.zeros(0, np.uint8),
                      np.zeros(1, np.int64), np.zeros(0, np.uint32), np.zeros(1, np.int64), np.zeros(8, np.uint8),
                      np.zeros(8, np.int8))
    engine.ref_upload(901, "empty", draft)
    rid = engine.readset_upload(901, empty)
    engine.batch_run([(901, 1, 3000, rid), (901, 2001, 2100, rid)], window=0, block=0, persistent=1,
                     copy_out=False, want_stats=False)
    got, _ = engine.batch_bed()
    ref = "".join(O.run_region(engine.config, draft, s, e, [empty], name="empty", vcf=O.BED)["bed"]
                  for s, e in [(1, 3000), (2001, 2100)])
    assert got == HEAD + ref
    assert "Gap" in ref
