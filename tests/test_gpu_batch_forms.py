"""The batched form's launch variants give the same bits: hipGraph replay
(several replays with other batches and host syncs in between, each checked
against the eager run), the autotuned window/block, and the persistent queue,
on a scaled many-contig workload checked against the oracle per contig."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests.helpers import assert_same_region, assert_same_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def many():
    from pilon_amd.engine import Context
    from pilon_amd.synth import SynthData, preset
    d = SynthData(preset("c3", scale=0.004))
    ctx = Context(default_config(), device=0)
    rids = []
    for ci in range(d.n):
        ctx.ref_upload(ci, d.names[ci], d.draft(ci))
        rids.append(ctx.readset_upload(ci, d.soa(ci)))
    regions = [(ci, 1, len(d.draft(ci)), rids[ci]) for ci in range(d.n)]
    yield d, ctx, regions
    ctx.close()


def _same(a, b, what):
    pa, ca, sa = a
    pb, cb, sb = b
    assert np.array_equal(pa, pb), what
    assert np.array_equal(ca.view(np.uint64), cb.view(np.uint64)), what
    assert sa == sb, what


def test_eager_matches_oracle(many):
    from oracle import oracle as O
    d, ctx, regions = many
    print(f"case: c3 x0.004 eager vs oracle, {len(regions)} regions", flush=True)
    pile, calls, stats = ctx.batch_run(regions, window=512, block=256, persistent=1)
    off = 0
    for (ci, s, e, _), st in zip(regions, stats):
        G = e - s + 1
        ref = O.run_region(ctx.config, d.draft(ci), s, e, [d.batch(ci)], name=d.names[ci])
        assert ref["status"] == 0
        assert_same_region(pile[:, off:off + G], calls[off:off + G], ref, f"contig {ci}")
        assert_same_stats(st, ref["stats"], f"contig {ci}")
        off += G


def test_graph_replays_match_eager(many):
    d, ctx, regions = many
    print("case: graph replays", flush=True)
    eager = ctx.batch_run(regions, window=512, block=256, persistent=1)
    assert ctx.batch_info()["window_ms"] > 0
    other = regions[: max(1, len(regions) // 3)]
    other_eager = ctx.batch_run(other, window=256, block=256, persistent=0)
    for i in range(4):
        got = ctx.batch_run(regions, window=512, block=256, persistent=1, use_graph=True)
        _same(got, eager, f"graph replay {i}")
        assert ctx.batch_info()["pipeline_ms"] > 0
        # the program's own eager work between replays
        _same(ctx.batch_run(other, window=256, block=256, persistent=0), other_eager, f"eager between replays {i}")
        got2 = ctx.batch_run(other, window=256, block=256, persistent=0, use_graph=True)
        _same(got2, other_eager, f"second graph replay {i}")


def test_repeated_passes_match(many):
    """pl_batch_repeat: 70 back-to-back replays of the captured batch (two
    rounds of queued passes) leave the same resident results as the eager
    run (its VCF, tracks and record counts); after an eager batch there is
    nothing to repeat."""
    from pilon_amd.engine import PilonError
    d, ctx, regions = many
    print("case: repeated graph passes", flush=True)
    eager = ctx.batch_run(regions, window=512, block=256, persistent=1)
    _same(ctx.batch_run(regions, window=512, block=256, persistent=1, use_graph=True), eager, "graph run")
    info0 = ctx.batch_info()
    vcf0, _ = ctx.batch_vcf()
    trk0, _ = ctx.batch_tracks()
    ctx.batch_repeat(70)
    info1 = ctx.batch_info()
    assert info1["pipeline_ms"] > 0
    assert (info1["n_records"], info1["n_candidates"], info1["n_deferred"]) == \
        (info0["n_records"], info0["n_candidates"], info0["n_deferred"])
    vcf1, _ = ctx.batch_vcf()
    trk1, _ = ctx.batch_tracks()
    assert vcf1 == vcf0
    assert np.array_equal(trk1, trk0)
    ctx.batch_repeat(0)
    other = regions[: max(1, len(regions) // 3)]
    ctx.batch_run(other, window=256, block=256, persistent=0)
    with pytest.raises(PilonError) as e:
        ctx.batch_repeat(1)
    assert e.value.status == 1


def test_autotuned_matches_fixed(many):
    d, ctx, regions = many
    print("case: autotuned window/block", flush=True)
    fixed = ctx.batch_run(regions, window=512, block=256, persistent=1)
    tuned = ctx.batch_run(regions, window=0, block=0, persistent=1)
    _same(tuned, fixed, "autotuned")
    info = ctx.batch_info()
    assert info["window"] in (256, 384, 512, 640, 768, 1024)
    # cached: the same shape does not search again and keeps the window
    again = ctx.batch_run(regions, window=0, block=0, persistent=1)
    _same(again, fixed, "autotuned again")
    assert ctx.batch_info()["window"] == info["window"]


@pytest.mark.parametrize("window,block", [(384, 256), (640, 512), (128, 1024)])
def test_window_shapes(many, window, block):
    d, ctx, regions = many
    print(f"case: window={window} block={block}", flush=True)
    ref = ctx.batch_run(regions, window=512, block=256, persistent=0)
    _same(ctx.batch_run(regions, window=window, block=block, persistent=1), ref, f"{window}/{block}")
