"""GPU parity: libpilonhip.so against the oracle, bit for bit (every PileUp field,
every BaseCall + disposition, the .changes text, the fixed bases, the region stats)."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from pilon_amd.reads import pack_reads
from tests.helpers import assert_same_region, assert_same_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


def _oracle(cfg, draft, s, e, batches, name):
    from oracle import oracle as O
    return O.run_region(cfg, draft, s, e, batches, name=name)


def _region(ctx, contig_id, s, e, batches=None, readset=None):
    from pilon_amd.engine import PileUpRegion
    pr = PileUpRegion(ctx, contig_id, s, e)
    if readset is not None:
        pr.useReadSet(readset)
    for b in batches or []:
        pr.addReads(b)
    pr.postProcess()
    return pr


def test_c1_full_region(engine):
    from pilon_amd.synth import SynthData, preset
    print("case: c1 full region", flush=True)
    d = SynthData(preset("c1"))
    draft, b = d.draft(0), d.batch(0)
    engine.ref_upload(0, d.names[0], draft)
    pr = _region(engine, 0, 1, len(draft), [b])
    ref = _oracle(engine.config, draft, 1, len(draft), [b], d.names[0])
    assert ref["status"] == 0
    assert_same_region(pr.pile, pr.calls, ref, "c1")
    assert_same_stats(pr.stats, ref["stats"], "c1")
    changes, bases = pr.fix()
    assert changes == ref["changes"]
    assert bases == ref["bases"]
    assert ref["stats"]["snps"] == 200


@pytest.mark.parametrize("window,block", [(256, 256), (512, 512), (896, 512), (1024, 1024), (1920, 256), (64, 512)])
@pytest.mark.parametrize("persistent", [0, 1])
def test_c2_small_batch_windows(engine, window, block, persistent):
    """Scaled c2 (clips, improper pairs, MAPQ-0 repeats, error indels, N bases)
    as three chunk regions plus the whole contig, in one batch."""
    from pilon_amd.synth import SynthData, preset
    print(f"case: c2 x0.04 window={window} block={block} persistent={persistent}", flush=True)
    p = preset("c2", scale=0.04)
    p["n_rate"] = 0.001
    p["err_indel_rate"] = 0.002
    d = SynthData(p)
    draft, b = d.draft(0), d.batch(0)
    engine.ref_upload(1, d.names[0], draft)
    rid = engine.readset_upload(1, b)
    L = len(draft)
    regions = [(1, L // 3), (L // 3 + 1, 2 * L // 3), (2 * L // 3 + 1, L), (5000, 5200), (1, L)]
    pile, calls, stats = engine.batch_run([(1, s, e, rid) for s, e in regions], window=window,
                                          persistent=persistent, block=block)
    off = 0
    for (s, e), st in zip(regions, stats):
        n = e - s + 1
        ref = _oracle(engine.config, draft, s, e, [b], d.names[0])
        assert ref["status"] == 0
        assert_same_region(pile[:, off:off + n], calls[off:off + n], ref, f"region {s}-{e}")
        assert_same_stats(st, ref["stats"], f"region {s}-{e}")
        off += n


def test_region_readset_and_fix(engine):
    from pilon_amd.synth import SynthData, preset
    print("case: c2 x0.02 region via resident read set + fix", flush=True)
    d = SynthData(preset("c2", scale=0.02))
    draft, b = d.draft(0), d.batch(0)
    engine.ref_upload(2, d.names[0], draft)
    rid = engine.readset_upload(2, b)
    L = len(draft)
    s, e = 30001, L
    pr = _region(engine, 2, s, e, readset=rid)
    ref = _oracle(engine.config, draft, s, e, [b], d.names[0])
    assert_same_region(pr.pile, pr.calls, ref, "readset region")
    changes, bases = pr.fix()
    assert changes == ref["changes"]
    assert bases == ref["bases"]


def test_two_bams_merge(engine):
    """Two add_reads calls (two BAMs) equal one call with all reads."""
    from pilon_amd.synth import SynthData, preset
    print("case: two add_reads batches", flush=True)
    d = SynthData(preset("c1", scale=0.3))
    draft, b = d.draft(0), d.batch(0)
    engine.ref_upload(3, d.names[0], draft)
    b1 = b.slice(0, b.n // 2)
    b2 = b.slice(b.n // 2, b.n)
    pr = _region(engine, 3, 1, len(draft), [b2, b1])
    ref = _oracle(engine.config, draft, 1, len(draft), [b], d.names[0])
    assert_same_region(pr.pile, pr.calls, ref, "two bams")


def test_empty_region(engine):
    print("case: region with no reads", flush=True)
    draft = b"ACGT" * 300
    engine.ref_upload(4, "empty", draft)
    pr = _region(engine, 4, 1, len(draft), [pack_reads([])])
    ref = _oracle(engine.config, draft, 1, len(draft), [pack_reads([])], "empty")
    assert_same_region(pr.pile, pr.calls, ref, "empty")
    assert pr.stats["read_count"] == 0


def test_deterministic(engine):
    from pilon_amd.synth import SynthData, preset
    print("case: determinism", flush=True)
    d = SynthData(preset("c2", scale=0.02, seed=77))
    draft, b = d.draft(0), d.batch(0)
    engine.ref_upload(5, d.names[0], draft)
    rid = engine.readset_upload(5, b)
    L = len(draft)
    out1 = engine.batch_run([(5, 1, L, rid)], window=512, persistent=1, block=256)
    out2 = engine.batch_run([(5, 1, L, rid)], window=1920, persistent=0, block=256)
    assert np.array_equal(out1[0], out2[0])
    assert np.array_equal(out1[1].view(np.uint64), out2[1].view(np.uint64))


def test_window_too_big_for_lds(engine):
    """A window/block pair that cannot fit a CU's LDS is refused, not launched."""
    from pilon_amd.engine import PilonError
    draft = b"ACGT" * 600
    engine.ref_upload(6, "lds", draft)
    rid = engine.readset_upload(6, pack_reads([]))
    with pytest.raises(PilonError) as e:
        engine.batch_run([(6, 1, len(draft), rid)], window=2048, block=1024)
    assert e.value.status == 1


@pytest.mark.parametrize("min_qual", [20, -5])
def test_min_qual_batched(min_qual):
    """PileUp.add's minQual test (PileUp.scala:77): with minQual 20 about a
    fifth of the synthetic bases fail it (no count, qualSum, mqSum or qSum);
    with -5 every base passes and the packed words take the signed path."""
    from oracle import oracle as O
    from pilon_amd.engine import Context
    from pilon_amd.synth import SynthData, preset
    print(f"case: minQual {min_qual}, c2 x0.04 batched", flush=True)
    p = preset("c2", scale=0.04)
    p.update({"err_indel_rate": 0.002})
    d = SynthData(p)
    cfg = default_config(min_qual=min_qual)
    draft, b = d.draft(0), d.batch(0)
    L = len(draft)
    ctx = Context(cfg, device=0)
    try:
        ctx.ref_upload(0, d.names[0], draft)
        rid = ctx.readset_upload(0, b)
        regions = [(1, L // 2), (L // 2 + 1, L)]
        pile, calls, stats = ctx.batch_run([(0, s, e, rid) for s, e in regions], window=512, block=256, persistent=1)
    finally:
        ctx.close()
    off = 0
    for k, (s, e) in enumerate(regions):
        ref = O.run_region(cfg, draft, s, e, [b], name=d.names[0])
        n = e - s + 1
        assert_same_region(pile[:, off:off + n], calls[off:off + n], ref, f"minQual {min_qual} {s}-{e}")
        assert_same_stats(stats[k], ref["stats"])
        off += n


def test_negative_quals(engine):
    """Quality bytes of 128 and more are negative Java bytes: with the
    default minQual 0 those bases fail PileUp.add's test while their
    neighbours pile (the read set takes the tested base loop)."""
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    print("case: negative quals", flush=True)
    d = SynthData(preset("c1"))
    draft, b = d.draft(0), d.batch(0)
    L = int(b.seq_off[-1])
    rng = np.random.default_rng(7)
    q = b.quals.copy()
    hit = rng.random(L) < 0.02
    q[:L][hit] = -rng.integers(1, 100, int(hit.sum()))
    b.quals = q
    engine.ref_upload(30, d.names[0], draft)
    rid = engine.readset_upload(30, b)
    n = len(draft)
    pile, calls, stats = engine.batch_run([(30, 1, n, rid)], window=512, block=256, persistent=1)
    ref = O.run_region(engine.config, draft, 1, n, [b], name=d.names[0])
    assert_same_region(pile, calls, ref, "negative quals")
    assert_same_stats(stats[0], ref["stats"])
