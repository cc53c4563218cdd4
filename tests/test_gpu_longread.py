"""Long-read BAMs (--nanopore, --pacbio) on the GPU against the oracle:
resident long-read sets through the batched form (pileup, calls, stats,
.changes and FASTA with excludeMotifs), and a region that mixes a frags BAM
with a nanopore BAM through the plain form (reads of both kinds merged,
each piled by its own BAM's rules)."""
import numpy as np
import pytest

from pilon_amd.abi import BAM_TYPES, default_config
from tests.helpers import assert_same_region, assert_same_stats

pytestmark = pytest.mark.gpu


def data():
    from pilon_amd.synth import SynthData, preset
    p = preset("c2", scale=0.04)
    p.update({"n_rate": 0.001, "err_indel_rate": 0.004, "n_indel": 40})
    return SynthData(p)


@pytest.mark.parametrize("lr", ["nanopore", "pacbio"])
def test_long_read_set_batched(lr):
    from oracle import oracle as O
    from pilon_amd.engine import Context, contig_regions
    print(f"case: resident {lr} set, batched", flush=True)
    d = data()
    cfg = default_config(longread=1)
    draft, b = d.draft(0), d.batch(0)
    L = len(draft)
    ctx = Context(cfg, device=0)
    try:
        ctx.ref_upload(0, d.names[0], draft)
        rid = ctx.readset_upload(0, b, bam_type=lr)
        regions = contig_regions(L, 60_000)
        pile, calls, stats = ctx.batch_run([(0, s, e, rid) for s, e in regions], persistent=1)
        changes, fasta = ctx.batch_fix()
    finally:
        ctx.close()
    want_c, want_f = O.process_regions_output(cfg, [(d.names[0], draft)], [(0, s, e) for s, e in regions],
                                              lambda k: [b], set_types=[BAM_TYPES[lr]])
    off = 0
    for k, (s, e) in enumerate(regions):
        ref = O.run_region(cfg, draft, s, e, [b], name=d.names[0], set_types=[BAM_TYPES[lr]])
        assert ref["status"] == 0
        n = e - s + 1
        assert_same_region(pile[:, off:off + n], calls[off:off + n], ref, f"{lr} {s}-{e}")
        assert_same_stats(stats[k], ref["stats"], f"{lr} {s}-{e}")
        off += n
    assert changes == want_c
    assert fasta == want_f
    # the long-read rules matter here: the frags reading of the same reads differs
    plain = O.run_region(cfg, draft, 1, regions[0][1], [b], name=d.names[0])
    assert not np.array_equal(plain["pile"], pile[:, :regions[0][1]])


def test_frags_and_nanopore_region():
    from oracle import oracle as O
    from pilon_amd.engine import Context, PileUpRegion
    print("case: region with a frags and a nanopore BAM", flush=True)
    d = data()
    cfg = default_config(longread=1)
    draft, b = d.draft(0), d.batch(0)
    L = len(draft)
    idx = np.arange(b.n)
    Fs, Ns = b.take(idx[idx % 3 != 0]), b.take(idx[idx % 3 == 0])
    ctx = Context(cfg, device=0)
    try:
        ctx.ref_upload(1, d.names[0], draft)
        for s, e in ((1, L), (1001, L - 999)):
            pr = PileUpRegion(ctx, 1, s, e)
            pr.addReads(Fs)
            pr.addReads(Ns, bam_type="nanopore")
            pr.postProcess()
            changes, bases = pr.fix()
            pr.close()
            ref = O.run_region(cfg, draft, s, e, [Fs, Ns], name=d.names[0], set_types=[0, 3])
            assert ref["status"] == 0
            assert_same_region(pr.pile, pr.calls, ref, f"mixed {s}-{e}")
            assert_same_stats(pr.stats, ref["stats"], f"mixed {s}-{e}")
            assert changes == ref["changes"] and bases == ref["bases"]
    finally:
        ctx.close()
