"""pl_readset_upload_many: read sets uploaded in one pass give the same
batched pileup, calls, stats, VCF and tracks as the same sets uploaded one
at a time (pl_readset_upload_typed); with empty sets, non-ACGT reads,
unpaired sets beside frags ones, and sets spread over several staging
buffers."""
import numpy as np
import pytest

from pilon_amd.abi import default_config

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


def run_both(engine, items, regions_of):
    one = [engine.readset_upload(c, r, t) for c, r, t in items]
    many = engine.readset_upload_many(items)
    assert len(many) == len(one) and len(set(many) | set(one)) == 2 * len(one)
    out = []
    for ids in (one, many):
        regions = regions_of(ids)
        pile, calls, stats = engine.batch_run(regions, window=0, block=0, persistent=1, copy_out=True,
                                              want_stats=True)
        vcf, _ = engine.batch_vcf()
        tracks, _ = engine.batch_tracks()
        out.append((pile.copy(), calls.copy(), stats, vcf, tracks.copy()))
    (p1, c1, s1, v1, t1), (p2, c2, s2, v2, t2) = out
    assert np.array_equal(p1, p2)
    assert np.array_equal(c1.view(np.uint8), c2.view(np.uint8))
    assert s1 == s2
    assert v1 == v2
    assert np.array_equal(t1, t2)
    return s1


@pytest.mark.parametrize("name,scale,extra", [("c3", 0.003, {"n_rate": 0.001}), ("c2", 0.02, {})])
def test_upload_many_matches_one_at_a_time(engine, name, scale, extra):
    from pilon_amd.reads import pack_reads
    from pilon_amd.synth import SynthData, preset
    print(f"case: upload many {name} x{scale}", flush=True)
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    base = 900 if name == "c3" else 950
    for ci in range(d.n):
        engine.ref_upload(base + ci, d.names[ci], d.draft(ci))
    items = [(base + ci, d.batch(ci), "frags") for ci in range(d.n)]
    # an empty set, and the first contig's reads again as an unpaired BAM
    items.append((base, pack_reads([]), "frags"))
    items.append((base, d.batch(0), "unpaired"))
    n = d.n

    def regions_of(ids):
        regs = [(base + ci, 1, len(d.draft(ci)), ids[ci]) for ci in range(n)]
        regs.append((base, 1, len(d.draft(0)), ids[n]))
        regs.append((base, 1, len(d.draft(0)), ids[n + 1]))
        return regs

    stats = run_both(engine, items, regions_of)
    assert sum(s["aligned_bases"] for s in stats) > 0
    assert stats[n]["read_count"] == 0


def test_upload_many_refusals(engine):
    from pilon_amd.engine import PilonError
    from pilon_amd.reads import pack_reads
    print("case: upload many refusals", flush=True)
    engine.ref_upload(990, "r", "ACGT" * 100)
    b = pack_reads([dict(pos=5, cigar="4M", seq="ACGT", qual=[30] * 4)])
    with pytest.raises(PilonError):
        engine.readset_upload_many([(990, b, "frags"), (991, b, "frags")])      # contig 991 not uploaded
    with pytest.raises(PilonError):
        engine.readset_upload_many([(990, b, "jumps")])
    with pytest.raises(PilonError):
        engine.readset_upload_many([(990, b, "nanopore")])                      # config.longread off
    assert engine.readset_upload_many([]) == []
