"""The stress and serving configurations at their full sizes against the oracle.

c5: 20 contigs of 25 Mb, 2x250 bp at 80x (about 164 M reads), diploid SNVs,
1,000 breakpoints with 30-200 bp clips, 50 hotspots of 10 kb at 2,000x.  The
60 regions GenomeFile.contigRegions cuts (three per contig) are piled in one
persistent, autotuned launch; every region's PileUp fields, calls and stats
are compared with the oracle, and the batch's .changes and polished FASTA
(offsets accumulated over each contig's chunks) with the oracle's.

c4: 50 contigs of 4 Mb at 50x resident; 1,000 polish requests of log-uniform
2 kb-1 Mb served with continuous batching: each request's pileup and calls
equal a run of that region alone, a sample equals the oracle, and every
request refused with PL_E_EDGE is one on which the oracle (Pilon) throws."""
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests.helpers import assert_same_region, assert_same_stats

pytestmark = pytest.mark.gpu


class _SoaOf:
    """A read set handed to the oracle straight from the generator's buffers."""

    def __init__(self, d, ci):
        self.d, self.ci = d, ci

    def soa(self):
        return self.d.soa(self.ci)


def log(msg):
    print(f"case: {msg}", flush=True)


def test_c5_full_size():
    from oracle import oracle as O
    from pilon_amd.engine import Context, contig_regions
    from pilon_amd.synth import SynthData, preset
    t0 = time.time()
    os.environ["PL_SYNTH_PROGRESS"] = "1"     # the generator reports its phases (minutes at this size)
    log("c5 full: generating 20 x 25 Mb, 2x250 bp at 80x")
    d = SynthData(preset("c5"))
    nreads = sum(d.n_reads(i) for i in range(d.n))
    log(f"c5 full: {d.n} contigs, {nreads} reads generated in {time.time() - t0:.0f}s")
    assert d.n == 20 and nreads > 150_000_000
    cfg = default_config()
    ctx = Context(cfg, device=0)
    try:
        regions = []
        for ci in range(d.n):
            ctx.ref_upload(ci, d.names[ci], d.draft(ci))
            rid = ctx.readset_upload(ci, d.soa(ci))
            for s, e in contig_regions(len(d.draft(ci))):
                regions.append((ci, s, e, rid))
            if ci % 5 == 4:
                log(f"c5 full: {ci + 1} contigs resident")
        assert len(regions) == 60
        assert [(s, e) for _, s, e, _ in regions[:3]] == [(1, 8333334), (8333335, 16666668), (16666669, 25000000)]
        log(f"c5 full: resident after {time.time() - t0:.0f}s; one persistent autotuned launch of 60 regions")
        pile, calls, stats = ctx.batch_run(regions, window=0, block=0, persistent=1)
        info = ctx.batch_info()
        log(f"c5 full: piled G={info['positions']} W={info['window']} B={info['block']} "
            f"{info['pipeline_ms']:.1f} ms; fix outputs")
        changes, fasta = ctx.batch_fix()
    finally:
        ctx.close()
    offs = np.concatenate([[0], np.cumsum([e - s + 1 for _, s, e, _ in regions])])
    drafts = {}

    def check_contig(ci):
        # a contig's chunks in order: its .changes offsets accumulate
        draft = d.draft(ci)
        nn = O.pilon_name(d.names[ci])
        offset, ch, seq, bc = 0, [], [], 0
        for k in range(3 * ci, 3 * ci + 3):
            _, s, e, _ = regions[k]
            ref = O.run_region(cfg, draft, s, e, [_SoaOf(d, ci)], name=d.names[ci], new_name=nn, offset=offset)
            assert ref["status"] == 0, k
            a, b = int(offs[k]), int(offs[k + 1])
            assert_same_region(pile[:, a:b], calls[a:b], ref, f"c5 region {k}")
            assert_same_stats(stats[k], ref["stats"], f"c5 region {k}")
            ch.append(ref["changes"])
            seq.append(ref["bases"].decode())
            offset += len(ref["bases"]) - (e - s + 1)
            bc += ref["stats"]["base_count"]
        return ci, "".join(ch), O.fasta_element(nn, "".join(seq)), bc

    want_c, want_f = [""] * d.n, [""] * d.n
    total = 0
    with ThreadPoolExecutor(16) as ex:
        for ci, ch, fa, bc in ex.map(check_contig, range(d.n)):
            want_c[ci], want_f[ci] = ch, fa
            total += bc
            log(f"c5 full: contig {ci} (3 regions) matches the oracle")
    assert total > 30_000_000_000
    assert changes == "".join(want_c)
    assert fasta == "".join(want_f)
    # the second and third chunks print with the offset of the earlier chunks' indels
    assert any(l.split()[0].split(":")[1].split("-")[0] != l.split()[1].split(":")[1].split("-")[0]
               for l in changes.splitlines())
    # hotspots: some column far deeper than the median
    depth = pile[0:4].sum(axis=0, dtype=np.int64)
    assert depth.max() > 10 * np.median(depth)
    log(f"c5 full: done in {time.time() - t0:.0f}s")


def test_c4_full_size_serving():
    from oracle import oracle as O
    from pilon_amd.engine import Context, Server
    from pilon_amd.synth import SynthData, preset, serve_requests
    t0 = time.time()
    os.environ["PL_SYNTH_PROGRESS"] = "1"
    log("c4 full: generating 50 x 4 Mb at 50x")
    d = SynthData(preset("c4"))
    assert d.n == 50 and sum(len(d.draft(i)) for i in range(d.n)) == 200_000_000
    cfg = default_config()
    ctx = Context(cfg, device=0)
    try:
        rids = []
        for ci in range(d.n):
            ctx.ref_upload(ci, d.names[ci], d.draft(ci))
            rids.append(ctx.readset_upload(ci, d.soa(ci)))
        reqs = serve_requests(d, 1000)
        log(f"c4 full: resident after {time.time() - t0:.0f}s; serving 1000 requests on 4 lanes")
        srv = Server(ctx, lanes=4, copy_out=True)
        mixed = {}
        try:
            tick = {srv.submit(ci, s, e, rids[ci], tag=k): k for k, (ci, s, e) in enumerate(reqs)}
            srv.drain()
            res = srv.poll(max_n=2000)
            assert len(res) == len(reqs)
            batched = [r for r in res if r["batch_size"] > 1]
            assert len(batched) > len(res) // 2        # requests did share batches
            for r in res:
                k = tick[r["ticket"]]
                if r["status"] == 0:
                    pile, calls = srv.fetch(r["ticket"])
                    mixed[k] = (pile, calls, r["stats"])
                else:
                    mixed[k] = (None, None, r["status"])
        finally:
            srv.close()
        log(f"c4 full: {len(res)} served, {sum(1 for v in mixed.values() if v[0] is None)} refused; alone vs mixed")
        # every request again alone (a batch of one region): identical results
        for k, (ci, s, e) in enumerate(reqs):
            pm, cm, sm = mixed[k]
            if pm is None:
                continue
            pa, ca, sa = ctx.batch_run([(ci, s, e, rids[ci])], window=512, block=256, persistent=1)
            assert np.array_equal(pa, pm) and np.array_equal(ca.view(np.uint64), cm.view(np.uint64)), k
            assert sa[0] == sm, k
            if k % 250 == 249:
                log(f"c4 full: {k + 1} requests identical alone and mixed")
    finally:
        ctx.close()
    # a sample against the oracle, and every refused request
    edge = [k for k, v in mixed.items() if v[0] is None]
    sample = list(range(0, len(reqs), 20))

    def check(k):
        ci, s, e = reqs[k]
        ref = O.run_region(cfg, d.draft(ci), s, e, [_SoaOf(d, ci)], name=d.names[ci])
        pm, cm, sm = mixed[k]
        if pm is None:
            assert sm == 4 and ref["status"] == 4, (k, sm, ref["status"], ref["edge_why"])
            return 0
        assert ref["status"] == 0, k
        assert_same_region(pm, cm, ref, f"c4 request {k}")
        assert_same_stats(sm, ref["stats"], f"c4 request {k}")
        return 1

    with ThreadPoolExecutor(16) as ex:
        n_ok = sum(ex.map(check, sorted(set(sample) | set(edge))))
    log(f"c4 full: {n_ok} sampled requests match the oracle, {len(edge)} refusals confirmed; "
        f"done in {time.time() - t0:.0f}s")
    assert n_ok >= 40
