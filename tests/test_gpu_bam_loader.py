"""Read sets decoded on the device from a BAM's record bytes (pl_loader_*
over libpilonbam's stream) against the same BAM decoded on the host
(pl_bam_open) and uploaded (pl_readset_upload_many): the batched pileup,
calls, stats, VCF and tracks of every reference are identical; on records
written literally from the SAMv1 specification (long CIGARs in the CG tag,
absent QUAL, non-ACGT codes, placed unmapped and unplaced records, a
reference without records) and on synthetic workloads in small slices."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests import bamspec as S
from tests.test_bam import known_records, w

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


def outputs(engine, regions):
    pile, calls, stats = engine.batch_run(regions, window=0, block=0, persistent=1, copy_out=True, want_stats=True)
    vcf, _ = engine.batch_vcf()
    tracks, _ = engine.batch_tracks()
    return pile.copy(), calls.copy(), stats, vcf, tracks.copy()


def both(engine, path, drafts, base, slice_bytes=1 << 20, skip=()):
    from pilon_amd.bam import BamFile
    n = len(drafts)
    for i, d in enumerate(drafts):
        engine.ref_upload(base + i, "ref%d" % i, d)
    with BamFile(path, threads=4) as b:
        keep = [i for i in range(n) if i not in skip]
        host = dict(zip(keep, engine.readset_upload_many([(base + i, b.reads(i), "frags") for i in keep])))
    dev = engine.readsets_from_bam(path, [-1 if i in skip else base + i for i in range(n)], threads=4,
                                   slice_bytes=slice_bytes)
    assert [dev[i] for i in skip] == [-1] * len(skip)
    got = []
    for ids in (host, {i: dev[i] for i in keep}):
        got.append(outputs(engine, [(base + i, 1, len(drafts[i]), ids[i]) for i in keep]))
    (p1, c1, s1, v1, t1), (p2, c2, s2, v2, t2) = got
    assert np.array_equal(p1, p2)
    assert np.array_equal(c1.view(np.uint8), c2.view(np.uint8))
    assert s1 == s2 and v1 == v2
    assert np.array_equal(t1, t2)
    return s1, engine.last_bam_stats


def test_spec_records(engine, tmp_path):
    print("case: loader on spec records", flush=True)
    rng = np.random.default_rng(5)
    drafts = ["".join("ACGT"[x] for x in rng.integers(0, 4, L)) for L in (1000, 500)]
    p = tmp_path / "known.bam"
    p.write_bytes(S.bgzf(known_records(), block=300))
    stats, bst = both(engine, p, drafts, 700, slice_bytes=1 << 17)
    assert stats[0]["read_count"] > 0 and bst["records"] == 6


def test_reference_without_records_and_skipped(engine, tmp_path):
    print("case: loader with an empty and a skipped reference", flush=True)
    rng = np.random.default_rng(6)
    drafts = ["".join("ACGT"[x] for x in rng.integers(0, 4, L)) for L in (400, 300, 500)]
    recs = [S.record(0, 10 * k, [w(30, "M")], drafts[0][10 * k:10 * k + 30], [30] * 30, name="a%d" % k)
            for k in range(20)]
    recs += [S.record(2, 7 * k, [w(25, "M")], drafts[2][7 * k:7 * k + 25], [25] * 25, name="c%d" % k)
             for k in range(30)]
    p = tmp_path / "gap.bam"
    p.write_bytes(S.bgzf(S.header([("r0", 400), ("r1", 300), ("r2", 500)]) + b"".join(recs), block=1000))
    stats, _ = both(engine, p, drafts, 720)
    assert [s["read_count"] for s in stats] == [20, 0, 30]
    both(engine, p, drafts, 730, skip=(0,))


@pytest.mark.parametrize("name,scale,extra", [("c2", 0.02, {"n_rate": 0.001}), ("c3", 0.003, {})])
def test_synthetic(engine, tmp_path, name, scale, extra):
    from pilon_amd.synth import SynthData, preset
    print(f"case: loader on {name} x{scale}", flush=True)
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    refs = [(d.names[i], len(d.draft(i))) for i in range(d.n)]
    path = tmp_path / f"{name}.bam"
    S.write_bam(path, refs, {i: d.batch(i) for i in range(d.n)})
    stats, bst = both(engine, path, [d.draft(i) for i in range(d.n)], 740 if name == "c2" else 800,
                      slice_bytes=1 << 20)
    assert sum(s["aligned_bases"] for s in stats) > 0
    if name == "c2":
        assert bst["records"] > 10000


def test_unsorted_and_out_of_order_refused(engine, tmp_path):
    from pilon_amd.engine import PilonError
    print("case: loader refusals", flush=True)
    engine.ref_upload(900, "x", "ACGT" * 50)
    engine.ref_upload(901, "y", "ACGT" * 50)
    h = S.header([("x", 200), ("y", 200)])
    unsorted = h + S.record(0, 50, [w(2, "M")], "AC", [9, 9]) + S.record(0, 10, [w(2, "M")], "AC", [9, 9])
    out_of_order = h + S.record(1, 5, [w(2, "M")], "AC", [9, 9]) + S.record(0, 10, [w(2, "M")], "AC", [9, 9])
    for name, blob in (("unsorted", unsorted), ("order", out_of_order)):
        p = tmp_path / f"{name}.bam"
        p.write_bytes(S.bgzf(blob))
        with pytest.raises(PilonError) as e:
            engine.readsets_from_bam(p, [900, 901])
        assert e.value.status == 1, name
