"""Parity at the configurations' real shapes: the full single-region c2 contig
(5 Mb at 100x: clips, improper pairs, MAPQ-0 repeats, error indels) and the
diploid c5 stress shape (2x250 bp, heterozygous SNVs, a 2000x hotspot,
30-200 bp breakpoint clips) at reduced genome size, every field, call,
changes line and fixed base against the oracle."""
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


def _check_contig(engine, d, ci, cid, regions, what):
    from oracle import oracle as O
    draft, b = d.draft(ci), d.batch(ci)
    engine.ref_upload(cid, d.names[ci], draft)
    rid = engine.readset_upload(cid, b)
    pile, calls, stats = engine.batch_run([(cid, s, e, rid) for s, e in regions], window=0, block=0, persistent=1)
    off = 0
    for (s, e), st in zip(regions, stats):
        n = e - s + 1
        ref = O.run_region(engine.config, draft, s, e, [b], name=d.names[ci])
        assert ref["status"] == 0, (what, s, e)
        assert_same_region(pile[:, off:off + n], calls[off:off + n], ref, f"{what} {s}-{e}")
        assert_same_stats(st, ref["stats"], f"{what} {s}-{e}")
        off += n
    return stats, pile


def test_c2_full_contig(engine):
    from pilon_amd.synth import SynthData, preset
    print("case: c2 full (5 Mb, 100x)", flush=True)
    d = SynthData(preset("c2"))
    L = len(d.draft(0))
    stats, _ = _check_contig(engine, d, 0, 10, [(1, L)], "c2")
    assert stats[0]["read_count"] > 3_000_000


def test_c2_full_fix_matches_oracle(engine):
    from oracle import oracle as O
    from pilon_amd.engine import PileUpRegion
    from pilon_amd.synth import SynthData, preset
    print("case: c2 full fix (changes + fixed bases)", flush=True)
    d = SynthData(preset("c2"))
    draft, b = d.draft(0), d.batch(0)
    engine.ref_upload(11, d.names[0], draft)
    rid = engine.readset_upload(11, b)
    pr = PileUpRegion(engine, 11, 1, len(draft))
    pr.useReadSet(rid)
    pr.postProcess()
    changes, bases = pr.fix()
    ref = O.run_region(engine.config, draft, 1, len(draft), [b], name=d.names[0])
    assert changes == ref["changes"]
    assert bases == ref["bases"]
    assert changes.count("\n") > 100


def test_c5_diploid_hotspot_shape(engine):
    from pilon_amd.synth import SynthData, preset
    print("case: c5 x0.02 (2x250, het SNVs, 2000x hotspot, breakpoint clips)", flush=True)
    d = SynthData(preset("c5", scale=0.02))
    L = len(d.draft(0))
    # three chunk regions like GenomeFile.contigRegions, plus the whole contig
    c = (L + 2) // 3
    regions = [(1, c), (c + 1, 2 * c), (2 * c + 1, L), (1, L)]
    stats, pile = _check_contig(engine, d, 0, 12, regions, "c5")
    assert stats[3]["read_count"] > 100_000
    # the hotspot is deep: some column holds far more than the mean depth
    depth = pile[0:4, -L:].sum(axis=0)
    assert depth.max() > 10 * np.median(depth)


class _SoaOf:
    """A read set handed to the oracle straight from the generator's buffers."""

    def __init__(self, d, ci):
        self.d, self.ci = d, ci

    def soa(self):
        return self.d.soa(self.ci)


def test_c3_full_batch(engine):
    """The many-contig batch at its full size (3,000 contigs, 60 Mb, 24 M
    reads at 60x), as the benchmark runs it (one persistent launch,
    autotuned shape): every field and call of every region against the
    oracle, the regions checked on 16 host threads."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    print("case: c3 full (3000 contigs, 60 Mb, 60x) batched vs oracle", flush=True)
    d = SynthData(preset("c3"))
    regions, rids = [], []
    for ci in range(d.n):
        engine.ref_upload(2000 + ci, d.names[ci], d.draft(ci))
        rid = engine.readset_upload(2000 + ci, d.soa(ci))
        regions.append((2000 + ci, 1, len(d.draft(ci)), rid))
    pile, calls, stats = engine.batch_run(regions, window=0, block=0, persistent=1)
    offs = np.concatenate([[0], np.cumsum([e - s + 1 for _, s, e, _ in regions])])
    print(f"case: c3 full piled, G={int(offs[-1])}; checking", flush=True)

    def check(k):
        cid, s, e, _ = regions[k]
        ci = cid - 2000
        ref = O.run_region(engine.config, d.draft(ci), s, e, [_SoaOf(d, ci)], name=d.names[ci],
                           new_name=O.pilon_name(d.names[ci]))
        assert ref["status"] == 0, k
        a, b = int(offs[k]), int(offs[k + 1])
        assert_same_region(pile[:, a:b], calls[a:b], ref, f"c3 region {k}")
        assert_same_stats(stats[k], ref["stats"], f"c3 region {k}")
        return ref["stats"]["base_count"], ref["changes"], O.fasta_element(O.pilon_name(d.names[ci]),
                                                                           ref["bases"].decode())

    with ThreadPoolExecutor(16) as ex:
        total = 0
        want_changes, want_fasta = [], []
        for k, (bc, ch, fa) in enumerate(ex.map(check, range(len(regions)))):
            total += bc
            want_changes.append(ch)
            want_fasta.append(fa)
            if k % 1000 == 999:
                print(f"case: c3 full, {k + 1} regions match", flush=True)
    assert total > 3_000_000_000
    # the whole batch's .changes file and polished FASTA (GenomeFile.processRegions' output loop)
    changes, fasta = engine.batch_fix()
    assert changes == "".join(want_changes)
    assert fasta == "".join(want_fasta)
    assert changes.count("\n") > 5000
