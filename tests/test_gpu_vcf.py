"""The VCF records formatted on the device (pl_batch_vcf) against the
oracle's, byte for byte: the known-answer cases (region API), synthetic
batches of many regions (clips, improper pairs, MAPQ-0 repeats, indels, N,
diploid hotspots), and the --diploid / --vcfqe variants."""
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


def first_diff(a, b):
    la, lb = a.splitlines(), b.splitlines()
    for k, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return k, x, y
    return min(len(la), len(lb)), len(la), len(lb)


@pytest.mark.parametrize("make", K.ALL, ids=[f.__name__ for f in K.ALL])
def test_vcf_kat(engine, make):
    from oracle import oracle as O
    from pilon_amd.engine import Context, PileUpRegion, PilonError
    case = make()
    print(f"case: vcf {make.__name__}", flush=True)
    cfg = default_config(**case.get("cfg", {}))
    own = Context(cfg, device=0) if case.get("cfg") else None
    ctx = own or engine
    try:
        ctx.ref_upload(150, "ctg", case["ref"].encode())
        pr = PileUpRegion(ctx, 150, case["start"], case["stop"])
        pr.addReads(K.batch_of(case))
        pr.postProcess()
        got, _ = pr.vcf()
        # the region's own reads: pl_batch_vcf refuses rather than read released memory
        with pytest.raises(PilonError) as e:
            ctx.batch_vcf()
        assert e.value.status == 1
        # after another batch on the context, the region's batch runs again
        ctx.ref_upload(151, "other", case["ref"].encode())
        pr2 = PileUpRegion(ctx, 151, case["start"], case["stop"])
        pr2.addReads(K.batch_of(case))
        pr2.postProcess()
        again, _ = pr.vcf()
        pr2.close()
        pr.close()
    finally:
        if own is not None:
            own.close()
    ref = O.run_region(cfg, case["ref"], case["start"], case["stop"], [K.batch_of(case)], name="ctg", vcf=O.VCF_ON)
    assert got == ref["vcf"], first_diff(got, ref["vcf"])
    assert again == got


@pytest.mark.parametrize("name,scale,extra,diploid,qe", [
    ("c1", 1.0, {}, False, False),
    ("c1", 1.0, {}, True, True),
    ("c2", 0.04, {"n_rate": 0.001, "err_indel_rate": 0.002}, False, False),
    ("c3", 0.003, {}, False, False),
    ("c5", 0.01, {}, True, False),
])
def test_vcf_synthetic_batch(engine, name, scale, extra, diploid, qe):
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    print(f"case: vcf {name} x{scale} diploid={diploid} qe={qe}", flush=True)
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    regions = []
    for ci in range(d.n):
        engine.ref_upload(400 + ci, d.names[ci], d.draft(ci))
        rid = engine.readset_upload(400 + ci, d.batch(ci))
        L = len(d.draft(ci))
        if name == "c5":   # chunk regions of a contig, as GenomeFile.contigRegions cuts them
            c = (L + 2) // 3
            regions += [(400 + ci, 1 + k * c, min(L, (k + 1) * c), rid) for k in range(3)]
        else:
            regions.append((400 + ci, 1, L, rid))
    engine.batch_run(regions, window=0, block=0, persistent=1, copy_out=False, want_stats=False)
    got, ms = engine.batch_vcf(diploid=diploid, vcfqe=qe)
    flags = O.VCF_ON | (O.VCF_DIPLOID if diploid else 0) | (O.VCF_QE if qe else 0)
    parts = []
    for cid, s, e, _ in regions:
        ci = cid - 400
        r = O.run_region(engine.config, d.draft(ci), s, e, [d.batch(ci)], name=d.names[ci], vcf=flags)
        assert r["status"] == 0
        parts.append(r["vcf"])
    ref = "".join(parts)
    assert got == ref, first_diff(got, ref)
    assert ms > 0


def test_vcf_long_names_and_empty_regions(engine):
    """A contig name of 400 characters makes each 256-locus chunk's text
    larger than the write pass's LDS stage (the direct-store path); a region
    with no reads prints LowCov records of empty pileups."""
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    print("case: vcf long contig name + empty region", flush=True)
    d = SynthData(preset("c1", scale=0.2))
    long_name = "contig_" + "x" * 393
    draft = d.draft(0)
    engine.ref_upload(500, long_name, draft)
    rid = engine.readset_upload(500, d.batch(0))
    from pilon_amd.reads import ReadBatch
    import numpy as np
    empty = ReadBatch(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.uint16), np.zeros(0, np.uint8),
                      np.zeros(1, np.int64), np.zeros(0, np.uint32), np.zeros(1, np.int64), np.zeros(8, np.uint8),
                      np.zeros(8, np.int8))
    engine.ref_upload(501, "empty", draft)
    rid_e = engine.readset_upload(501, empty)
    L = len(draft)
    regions = [(500, 1, L, rid), (501, 1, 3000, rid_e), (500, 5000, L, rid)]
    engine.batch_run(regions, window=0, block=0, persistent=1, copy_out=False, want_stats=False)
    got, _ = engine.batch_vcf()
    ref = "".join([
        O.run_region(engine.config, draft, 1, L, [d.batch(0)], name=long_name, vcf=O.VCF_ON)["vcf"],
        O.run_region(engine.config, draft, 1, 3000, [empty], name="empty", vcf=O.VCF_ON)["vcf"],
        O.run_region(engine.config, draft, 5000, L, [d.batch(0)], name=long_name, vcf=O.VCF_ON)["vcf"],
    ])
    assert got == ref, first_diff(got, ref)
    assert max(len(l) for l in got.splitlines()) > 400


def test_vcf_duplication_record(engine):
    """The two-BAM case whose second BAM doubles 15 kb (copy number 2 over
    more than 10,000 loci): the device's VCF holds the oracle's writeDup
    line before the records of the event's first locus, byte for byte."""
    from oracle import oracle as O
    from pilon_amd.engine import PileUpRegion
    from tests.test_oracle_bed import dup_case
    print("case: vcf duplication record, two BAMs", flush=True)
    d, draft, sets = dup_case()
    engine.ref_upload(902, d.names[0], draft)
    pr = PileUpRegion(engine, 902, 1, len(draft))
    for b in sets:
        pr.addReads(b)
    pr.postProcess()
    got, _ = pr.vcf()
    pr.close()
    ref = O.run_region(engine.config, draft, 1, len(draft), sets, name=d.names[0], vcf=O.VCF_ON)["vcf"]
    assert got == ref, first_diff(got, ref)
    assert got.count("<DUP>") == 1


def test_vcf_stream_equals_text(engine):
    """pl_batch_vcf_stream hands over exactly pl_batch_vcf's text, in pieces
    of at most the staging size, for a batch larger than one piece; a sink
    that stops ends the call with PL_E_INVAL."""
    from pilon_amd.engine import PilonError
    from pilon_amd.synth import SynthData, preset
    print("case: vcf stream vs text (c3 x0.05)", flush=True)
    d = SynthData(preset("c3", scale=0.05))
    regions = []
    for ci in range(d.n):
        engine.ref_upload(600 + ci, d.names[ci], d.draft(ci))
        regions.append((600 + ci, 1, len(d.draft(ci)), engine.readset_upload(600 + ci, d.soa(ci))))
    engine.batch_run(regions, window=0, block=0, persistent=1, copy_out=False, want_stats=False)
    text, _ = engine.batch_vcf()
    pieces = []
    n, ms = engine.batch_vcf_stream(pieces.append)
    assert n == len(text) and b"".join(pieces).decode() == text
    assert len(pieces) > 2 and max(len(p) for p in pieces) <= 32 << 20 and ms > 0
    with pytest.raises(PilonError) as e:
        engine.batch_vcf_stream(lambda b: False)
    assert e.value.status == 1
