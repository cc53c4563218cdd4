"""pl_batch_run_noalloc through the C ABI against pl_batch_run: the same
pileup, calls and stats bit for bit on two calls in a row (the second with
a smaller batch, so it runs over what the first left in the scratch array),
and the refusal of a batch beyond the array's position limit."""
import ctypes as C

import numpy as np
import pytest

from pilon_amd.abi import NFIELDS, PlBatchOpts, PlRegionDesc, PlRegionStats, call_dtype, default_config

pytestmark = pytest.mark.gpu


def _run(fn, ctx, regions, opts):
    n = len(regions)
    descs = (PlRegionDesc * n)(*[PlRegionDesc(c, s, e, r) for c, s, e, r in regions])
    G = sum(e - s + 1 for _, s, e, _ in regions)
    pile = np.zeros((NFIELDS, G), np.int32)
    calls = np.zeros(G, call_dtype())
    stats = (PlRegionStats * n)()
    st = fn(ctx.h, descs, n, C.byref(opts), pile.ctypes.data, calls.ctypes.data, stats)
    return st, pile, calls, [stats[i].as_dict() for i in range(n)]


def test_noalloc_matches_batch_run():
    from pilon_amd.engine import Context, lib
    from pilon_amd.synth import SynthData, preset
    print("case: no-allocation batch vs pl_batch_run, two calls", flush=True)
    p = preset("c5", scale=0.04)            # 1 Mb contig, 2x250 bp at 80x, het SNVs, hotspot, breakpoints
    d = SynthData(p)
    ctx = Context(default_config(), device=0)
    try:
        draft = d.draft(0)
        L = len(draft)
        ctx.ref_upload(0, d.names[0], draft)
        rid = ctx.readset_upload(0, d.soa(0))
        opts = PlBatchOpts(window=512, block=256, persistent=1, use_graph=0, copy_out=1)
        # first call: 4 regions, 1,000,000 positions; second: 2 regions, 120,000
        big = [(0, 1, 250_000, rid), (0, 250_001, 600_000, rid), (0, 600_001, 900_000, rid), (0, 900_001, L, rid)]
        small = [(0, 10_001, 70_000, rid), (0, 400_001, 460_000, rid)]
        assert sum(e - s + 1 for _, s, e, _ in big) <= 1 << 20
        for k, regions in enumerate([big, small]):
            st, pile, calls, stats = _run(lib().pl_batch_run_noalloc, ctx, regions, opts)
            assert st == 0, lib().pl_last_error()
            st, pile2, calls2, stats2 = _run(lib().pl_batch_run, ctx, regions, opts)
            assert st == 0, lib().pl_last_error()
            assert np.array_equal(pile, pile2), k
            assert np.array_equal(calls.view(np.uint64), calls2.view(np.uint64)), k
            assert stats == stats2, k
            assert stats[0]["read_count"] > 0
            print(f"case: no-allocation call {k + 1} matches ({len(regions)} regions)", flush=True)
        # beyond 2^20 positions: refused before anything runs
        ctx.ref_upload(1, "long", draft * 2)
        rid2 = ctx.readset_upload(1, d.soa(0))
        st, _, _, _ = _run(lib().pl_batch_run_noalloc, ctx, [(1, 1, 2 * L, rid2)], opts)
        assert st == 1 and b"no-allocation limit" in lib().pl_last_error()
    finally:
        ctx.close()
