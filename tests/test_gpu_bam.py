"""BAM -> engine: the read batches the BAM reader decodes go through the GPU
path (pl_readset_upload and pl_region_add_reads) and every region matches the
oracle run on the generator's own batches."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests.helpers import assert_same_region, assert_same_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


@pytest.mark.parametrize("name,scale,extra", [("c2", 0.02, {"n_rate": 0.001}), ("c3", 0.002, {})])
def test_bam_batches_on_gpu_match_oracle(engine, tmp_path, name, scale, extra):
    from oracle import oracle as O
    from pilon_amd.bam import BamFile
    from pilon_amd.synth import SynthData, preset
    print(f"case: {name} x{scale} via BAM", flush=True)
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    path = tmp_path / f"{name}.bam"
    d.write_bam(path)
    with BamFile(path, threads=4) as b:
        assert b.names == d.names
        regions = []
        for ci in range(d.n):
            engine.ref_upload(200 + ci, b.names[ci], d.draft(ci))
            rid = engine.readset_upload(200 + ci, b.reads(ci))
            regions.append((200 + ci, 1, b.lens[ci], rid))
        pile, calls, stats = engine.batch_run(regions, window=0, block=0, persistent=1)
    off = 0
    for ci, (_, s, e, _) in enumerate(regions):
        n = e - s + 1
        ref = O.run_region(engine.config, d.draft(ci), s, e, [d.batch(ci)], name=d.names[ci])
        assert ref["status"] == 0
        assert_same_region(pile[:, off:off + n], calls[off:off + n], ref, f"{name} contig {ci}")
        assert_same_stats(stats[ci], ref["stats"], f"{name} contig {ci}")
        off += n


def test_bam_region_add_reads(engine, tmp_path):
    """PileUpRegion.addRead fed from the BAM reader: changes and fixed bases."""
    from oracle import oracle as O
    from pilon_amd.bam import BamFile
    from pilon_amd.engine import PileUpRegion
    from pilon_amd.synth import SynthData, preset
    print("case: c1 via BAM, region API", flush=True)
    d = SynthData(preset("c1"))
    path = tmp_path / "c1.bam"
    d.write_bam(path)
    with BamFile(path) as b:
        engine.ref_upload(300, b.names[0], d.draft(0))
        L = b.lens[0]
        pr = PileUpRegion(engine, 300, 1, L)
        pr.addReads(b.reads(0))
        pr.postProcess()
        changes, bases = pr.fix()
        pr.close()
    ref = O.run_region(engine.config, d.draft(0), 1, L, [d.batch(0)], name=d.names[0])
    assert changes == ref["changes"] and bases == ref["bases"]
    assert changes.count("\n") > 50


def test_bam_region_through_index(engine, tmp_path):
    """BamFile.process's queryOverlapping through the BAI (BamFile.scala:
    118-119): each region piled from its index query of [start - 10 kb,
    stop + 10 kb] equals the same region piled from the whole file, field
    for field, and the oracle."""
    from oracle import oracle as O
    from pilon_amd.bam import BamFile, IndexedBamFile
    from pilon_amd.engine import PileUpRegion
    from pilon_amd.synth import SynthData, preset
    from tests import bamspec as S
    from tests.helpers import assert_same_region, assert_same_stats
    print("case: c2 x0.05 regions through the BAI index", flush=True)
    d = SynthData(preset("c2", scale=0.05))
    draft, b = d.draft(0), d.batch(0)
    L = len(draft)
    path = tmp_path / "c2.bam"
    S.write_bam_indexed(str(path), [(d.names[0], L)], {0: b})
    engine.ref_upload(310, d.names[0], draft)
    with BamFile(path) as whole, IndexedBamFile(path) as ib:
        for s, e in [(1, 40_000), (100_001, 160_000), (L - 30_000, L)]:
            q = ib.query(0, max(s - 10000, 0), min(e + 10000, L))
            out = []
            for reads in (q, whole.reads(0)):
                pr = PileUpRegion(engine, 310, s, e)
                pr.addReads(reads)
                pr.postProcess()
                out.append((pr.pile, pr.calls, pr.stats, pr.fix()))
                pr.close()
            (p1, c1, s1, f1), (p2, c2, s2, f2) = out
            assert np.array_equal(p1, p2) and np.array_equal(c1.view(np.uint64), c2.view(np.uint64)), (s, e)
            assert s1 == s2 and f1 == f2, (s, e)
            ref = O.run_region(engine.config, draft, s, e, [b], name=d.names[0])
            assert_same_region(p1, c1, ref, f"indexed {s}-{e}")
            assert_same_stats(s1, ref["stats"])
            assert q.n < b.n
