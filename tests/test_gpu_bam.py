"""BAM -> engine: the read batches the BAM reader decodes go through the GPU
path (pl_readset_upload and pl_region_add_reads) and every region matches the
oracle run on the generator's own batches."""
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
