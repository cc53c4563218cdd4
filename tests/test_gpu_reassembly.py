"""Region-level local reassembly on the device path (pl_region_reassemble:
gaps and possible breaks from the device's issue regions, recruitment from
the read sets, the K-mer graph on the GPU, the walk and joins on the host,
the fixes in pl_region_fix and pl_region_bed) against the oracle's
restatement of identifyAndFixIssues, exactly: .changes, polished bases,
the VCF with its writeFixRecord records, Pilon.bed lines before and after
the VCF, and logFix lines."""
import pytest

from pilon_amd.abi import default_config

from . import asm_cases as AC

pytestmark = pytest.mark.gpu


def cfg(**kw):
    return default_config(fix_snps=1, fix_indels=1, **kw)


def both(c, conf, sets_of=None, **kw):
    from oracle import asm as A
    from pilon_amd.engine import Context, PileUpRegion
    d = c["draft"]
    sets = sets_of(c) if sets_of else [(c["reads"], "frags", 500)]
    ref = A.reassemble_region(conf, d, 1, len(d), sets, name="ctg", **kw)
    ref_v = A.reassemble_region(conf, d, 1, len(d), sets, name="ctg", vcf=True, **kw)
    assert ref["status"] == 0 and ref_v["status"] == 0
    ctx = Context(conf, device=0)
    try:
        ctx.ref_upload(0, "ctg", d)
        reg = PileUpRegion(ctx, 0, 1, len(d))
        for b, t, _ in sets:
            reg.addReads(b, bam_type=t)
        reg.postProcess()
        opts = {k: int(v) for k, v in kw.items()}
        n, log = reg.reassemble(sets, **opts)
        changes, bases = reg.fix()
        bed, _ = reg.bed(track_line=False)
        vcf, _ = reg.vcf()
        bed_after, _ = reg.bed(track_line=False)
        reg.close()
    finally:
        ctx.close()
    assert log == ref["log"]
    assert changes == ref["changes"]
    assert bases.decode() == ref["bases"]
    assert bed == ref["bed"]
    assert vcf == ref_v["vcf"]
    assert bed_after == ref_v["bed"]
    return ref_v, n


def test_planted_region():
    print("case: planted region seed 31", flush=True)
    c = AC.planted_region()
    ref, n = both(c, cfg())
    assert n == 2 and ref["bases"] == c["truth"]


@pytest.mark.parametrize("kw", [dict(fix_gaps=False), dict(fix_local=False), dict(tr_safe=False, multi_closure=True),
                                dict(break_radius=300), dict(min_gap=5, gap_margin=10)],
                         ids=lambda kw: "-".join("%s%s" % kv for kv in sorted(kw.items())))
def test_planted_region_options(kw):
    print("case: planted region options", kw, flush=True)
    both(AC.planted_region(seed=34), cfg(), **kw)


def test_several_attempts():
    print("case: two gaps and a deletion", flush=True)
    c = AC.planted_region(seed=33, n=20000, gap=(4000, 120), deletion=(15000, 25), snps=(1000, 18000))
    d = c["draft"]
    c["draft"] = d[:9000] + "N" * 60 + d[9060:]
    ref, n = both(c, cfg())
    assert n == 3 and ref["bases"] == c["truth"]


def test_sparse_reads_and_a_tandem_repeat():
    """Reads every 25 bases (depth 6: some K-mers fall below minDepth) and a
    gap holding a tandem repeat: partial fills, opened gaps and logged
    NoSolution breaks, whatever Pilon makes of them, the same here."""
    print("case: sparse reads", flush=True)
    both(AC.planted_region(seed=35, step=25), cfg())
    print("case: tandem repeat in a gap", flush=True)
    t = AC.planted_region(seed=36)["truth"]
    both(AC.planted_region(seed=36, truth=t[:3020] + t[3020:3032] * 7 + t[3104:]), cfg())


def test_two_read_sets_unpaired_and_iupac():
    """A frags set and an unpaired set (both recruited), IUPAC codes for
    ambiguous bases."""
    print("case: frags + unpaired, iupac", flush=True)
    c = AC.planted_region(seed=37)
    half = c["reads"]
    idx_a = list(range(0, half.n, 2))
    idx_b = list(range(1, half.n, 2))
    a, b = half.take(idx_a), half.take(idx_b)
    both(c, cfg(iupac=1), sets_of=lambda c: [(a, "frags", 500), (b, "unpaired", 300)])


def test_jumps_set_is_refused():
    from pilon_amd.engine import Context, PileUpRegion, PilonError
    print("case: jumps set refused", flush=True)
    c = AC.planted_region(seed=38)
    d = c["draft"]
    ctx = Context(cfg(), device=0)
    try:
        ctx.ref_upload(0, "ctg", d)
        reg = PileUpRegion(ctx, 0, 1, len(d))
        reg.addReads(c["reads"])
        reg.postProcess()
        with pytest.raises(PilonError):
            reg.reassemble([(c["reads"], "jumps", 10000)])
        reg.close()
    finally:
        ctx.close()
