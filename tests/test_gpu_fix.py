"""The fix outputs of the batched and served forms against the oracle, byte
for byte: pl_batch_fix (GenomeFile.processRegions' .changes lines with the
per-element newName and accumulated offsets, and the FASTA of the fixed
chunks concatenated) on a chunked contig whose later chunks carry a
nonzero offset, on --targets elements (duplicates, a name holding '|'),
and pl_serve_fetch_fix of served requests."""
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


def first_diff(a, b):
    la, lb = a.splitlines(), b.splitlines()
    for k, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return k, x, y
    return min(len(la), len(lb)), len(la), len(lb)


def test_batch_fix_chunked_contig(engine):
    """A 500 kb c5-shaped contig (diploid SNVs, hotspot, breakpoints, error
    indels) named with a '|', cut by contigRegions at chunk size 150,000 into
    4 regions: chunks 2-4 print their .changes with the offset of the
    indels fixed in the chunks before them, and the FASTA is one element."""
    from oracle import oracle as O
    from pilon_amd.engine import contig_regions
    from pilon_amd.synth import SynthData, preset
    print("case: batch fix, c5 x0.02 in 150 kb contigRegions chunks", flush=True)
    p = preset("c5", scale=0.02)
    p.update({"err_indel_rate": 0.0005, "n_indel": 60})
    d = SynthData(p)
    draft, b = d.draft(0), d.batch(0)
    name = "gi|77|c5"
    engine.ref_upload(300, name, draft)
    rid = engine.readset_upload(300, b)
    chunks = contig_regions(len(draft), 150_000)
    assert chunks == O.contig_regions(len(draft), 150_000) and len(chunks) == 4
    engine.batch_run([(300, s, e, rid) for s, e in chunks], persistent=1, copy_out=False)
    changes, fasta = engine.batch_fix()
    want_c, want_f = O.process_regions_output(engine.config, [(name, draft)], [(0, s, e) for s, e in chunks],
                                              lambda k: [b])
    assert changes == want_c, first_diff(changes, want_c)
    assert fasta == want_f, first_diff(fasta, want_f)
    assert fasta.startswith(">gi|77|c5|pilon\n")
    # a later chunk's lines carry the offset of the earlier chunks' indels
    later = [l for l in changes.splitlines() if int(l.split()[0].split(":")[1].split("-")[0]) > chunks[0][1]]
    assert later and any(l.split()[0].split(":")[1] != l.split()[1].split(":")[1] for l in later)


def test_batch_fix_target_elements(engine):
    """--targets regions of two contigs, one listed twice: each target is its
    own element (its own newName, offset 0, FASTA element); the element
    argument groups regions as GenomeFile.regions does."""
    from oracle import oracle as O
    from pilon_amd.engine import parse_targets
    from pilon_amd.synth import SynthData, preset
    print("case: batch fix, targets of c1-shaped contigs", flush=True)
    p = preset("c1")
    p.update({"lens": [100_000, 80_000]})
    d = SynthData(p)
    contigs = [(d.names[i], d.draft(i)) for i in range(2)]
    rids = []
    for i, (nm, dr) in enumerate(contigs):
        engine.ref_upload(310 + i, nm, dr)
        rids.append(engine.readset_upload(310 + i, d.batch(i)))
    L0, L1 = len(contigs[0][1]), len(contigs[1][1])
    spec = f"{contigs[0][0]}:1001-60000,{contigs[1][0]},{contigs[0][0]}:1001-60000,{contigs[0][0]}:50000-{L0}"
    targets = parse_targets(spec, [(nm, len(dr)) for nm, dr in contigs])
    assert targets == O.parse_target_string(spec, [(nm, len(dr)) for nm, dr in contigs])
    engine.batch_run([(310 + c, s, e, rids[c]) for c, s, e in targets], persistent=1, copy_out=False)
    changes, fasta = engine.batch_fix(element=list(range(len(targets))))
    want_c, want_f = O.process_regions_output(engine.config, contigs, targets, lambda k: [d.batch(targets[k][0])],
                                              element=list(range(len(targets))))
    assert changes == want_c, first_diff(changes, want_c)
    assert fasta == want_f, first_diff(fasta, want_f)
    assert fasta.count(">") == 4 and L1 > 0
    # without element ids the two adjacent targets of contig 0 at the end
    # would be one element: the grouping is the caller's
    _, f2 = engine.batch_fix(want_changes=False)
    assert f2.count(">") == 3


def test_served_fix_outputs(engine):
    """Served requests kept with PL_SERVE_KEEP_FIX: each request's .changes
    and fixed bases equal the oracle's for its region, with the default
    polished name and with a caller's name and offset; the pileup of the
    same ticket is still fetchable afterwards."""
    from oracle import oracle as O
    from pilon_amd.engine import Server
    from pilon_amd.synth import SynthData, preset
    print("case: served fix outputs", flush=True)
    p = preset("c2", scale=0.05)
    p.update({"err_indel_rate": 0.001, "n_indel": 40})
    d = SynthData(p)
    draft, b = d.draft(0), d.batch(0)
    engine.ref_upload(320, d.names[0], draft)
    rid = engine.readset_upload(320, b)
    rng = np.random.default_rng(4)
    L = len(draft)
    reqs = []
    for k in range(24):
        ln = int(rng.integers(2000, 60000))
        s = int(rng.integers(1, L - ln))
        reqs.append((s, s + ln - 1))
    srv = Server(engine, lanes=2, copy_out=True, keep_fix=True)
    try:
        tickets = {srv.submit(320, s, e, rid, tag=k): k for k, (s, e) in enumerate(reqs)}
        srv.drain()
        res = srv.poll(max_n=100)
        assert len(res) == len(reqs)
        n_ok = 0
        for r in res:
            k = tickets[r["ticket"]]
            s, e = reqs[k]
            if r["status"] != 0:      # Pilon throws on this region: the oracle must too
                assert r["status"] == 4, (k, r["status"])
                assert O.run_region(engine.config, draft, s, e, [b], name=d.names[0])["status"] == 4, k
                continue
            n_ok += 1
            nn, off = (None, 0) if k % 2 == 0 else ("renamed", 1000 + k)
            ch, bases = srv.fetch_fix(r["ticket"], new_name=nn, offset=off)
            ref = O.run_region(engine.config, draft, s, e, [b], name=d.names[0],
                               new_name=nn if nn else O.pilon_name(d.names[0]), offset=off)
            assert ch == ref["changes"], (k, first_diff(ch, ref["changes"]))
            assert bases == ref["bases"], k
            pile, _ = srv.fetch(r["ticket"])
            assert np.array_equal(pile, ref["pile"]), k
        assert n_ok >= 20
    finally:
        srv.close()
