"""A region given a frags BAM and a jumps BAM (pl_region_add_reads_typed)
against the oracle: processBam (GenomeRegion.scala:287-300) leaves jumps
BAMs out of fragment coverage, so the delta, dip and copy-number tracks,
Pilon.bed and the VCF (its copy numbers) follow the frags reads alone while
the pileup and calls hold both."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests.test_oracle_tracks import jumps_split

pytestmark = pytest.mark.gpu

HEAD = 'track description="Issues found by Pilon" name="Pilon"\n'


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


@pytest.mark.parametrize("name,scale,extra", [
    ("c1", 1.0, {}),
    ("c2", 0.04, {"n_rate": 0.001, "err_indel_rate": 0.002}),
])
def test_jumps_region(engine, name, scale, extra):
    from oracle import oracle as O
    from pilon_amd.engine import PileUpRegion
    from pilon_amd.synth import SynthData, preset
    print(f"case: jumps region {name} x{scale}", flush=True)
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    draft, L = d.draft(0), len(d.draft(0))
    b, Fs, Js = jumps_split(d)
    engine.ref_upload(190, d.names[0], draft)
    pr = PileUpRegion(engine, 190, 1, L)
    pr.addReads(Fs)
    pr.addReads(Js, bam_type="jumps")
    pr.postProcess()
    tracks, _ = pr.tracks()
    bed, _ = pr.bed()
    vcf, _ = pr.vcf()
    # after another region on the context the region's batches run again
    engine.ref_upload(191, "other", draft)
    pr2 = PileUpRegion(engine, 191, 1, L)
    pr2.addReads(b)
    pr2.postProcess()
    plain, _ = pr2.tracks()
    again, _ = pr.tracks()
    bed2, _ = pr.bed()
    pr2.close()
    pr.close()
    ref = O.run_region(engine.config, draft, 1, L, [Fs, Js], name=d.names[0], set_types=[0, 1],
                       vcf=O.TRACKS | O.BED | O.VCF_ON)
    assert np.array_equal(pr.pile, ref["pile"])
    for t in range(len(O.TRACK_NAMES)):
        bad = np.nonzero(tracks[t] != ref["tracks"][t])[0]
        assert bad.size == 0, (O.TRACK_NAMES[t], bad[:5])
    assert bed == HEAD + ref["bed"]
    assert vcf == ref["vcf"]
    assert np.array_equal(again, tracks) and bed2 == bed
    # the same reads as one frags BAM: fragment tracks differ
    full = O.run_region(engine.config, draft, 1, L, [b], name="other", vcf=O.TRACKS)
    assert np.array_equal(plain, full["tracks"])
    assert not np.array_equal(plain[6], tracks[6])
