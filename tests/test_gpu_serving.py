"""The serving form: requests admitted by estimated device memory and run with
continuous batching on several lanes.  Every result is the same bits as the
region run alone; the admission limits answer as documented."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests.helpers import assert_same_region, assert_same_stats

pytestmark = pytest.mark.gpu

PL_E_BUSY, PL_E_REJECTED = 6, 7


@pytest.fixture(scope="module")
def resident():
    from pilon_amd.engine import Context
    from pilon_amd.synth import SynthData, preset
    d = SynthData(preset("c4", scale=0.05))
    ctx = Context(default_config(), device=0)
    rids = []
    for ci in range(d.n):
        ctx.ref_upload(ci, d.names[ci], d.draft(ci))
        rids.append(ctx.readset_upload(ci, d.soa(ci)))
    yield d, ctx, rids
    ctx.close()


def _wait_all(srv, n, seconds=60):
    import time
    got = []
    t_end = time.time() + seconds
    while len(got) < n:
        assert time.time() < t_end, f"only {len(got)} of {n} requests finished"
        got += srv.poll(wait_ms=1000)
    return got


def test_alone_equals_mixed(resident):
    from pilon_amd.engine import Server
    from pilon_amd.synth import serve_requests
    d, ctx, rids = resident
    reqs = serve_requests(d, 48, 2_000, 150_000, seed=11)
    print(f"case: serving alone vs mixed, {len(reqs)} requests", flush=True)
    alone = []
    for ci, s, e in reqs:
        pile, calls, st = ctx.batch_run([(ci, s, e, rids[ci])], window=512, block=256, persistent=0)
        alone.append((pile, calls, st[0]))
    srv = Server(ctx, lanes=3, max_batch=8, copy_out=True)
    try:
        for i, (ci, s, e) in enumerate(reqs):
            srv.submit(ci, s, e, rids[ci], tag=i)
        res = _wait_all(srv, len(reqs))
        assert sorted(r["tag"] for r in res) == list(range(len(reqs)))
        assert max(r["batch_size"] for r in res) > 1, "no request shared a batch"
        for r in res:
            assert r["status"] == 0, r
            pile, calls = srv.fetch(r["ticket"])
            ap, ac, ast = alone[r["tag"]]
            assert np.array_equal(pile, ap), r["tag"]
            assert np.array_equal(calls.view(np.uint64), ac.view(np.uint64)), r["tag"]
            assert r["stats"] == ast, r["tag"]
            assert r["latency_ms"] >= r["queue_ms"] >= 0
    finally:
        srv.close()


def test_served_matches_oracle(resident):
    from oracle import oracle as O
    from pilon_amd.engine import Server
    d, ctx, rids = resident
    reqs = [(0, 1, 40_000), (1, 150_001, len(d.draft(1))), (2, 90_000, 90_999)]
    print("case: served vs oracle", flush=True)
    srv = Server(ctx, lanes=2, max_batch=4, copy_out=True)
    try:
        for i, (ci, s, e) in enumerate(reqs):
            srv.submit(ci, s, e, rids[ci], tag=i)
        for r in _wait_all(srv, len(reqs)):
            ci, s, e = reqs[r["tag"]]
            pile, calls = srv.fetch(r["ticket"])
            ref = O.run_region(ctx.config, d.draft(ci), s, e, [d.batch(ci)], name=d.names[ci])
            assert ref["status"] == 0
            assert_same_region(pile, calls, ref, f"request {r['tag']}")
            assert_same_stats(r["stats"], ref["stats"], f"request {r['tag']}")
    finally:
        srv.close()


def test_admission_budget_and_limits(resident):
    """Estimates: ~80 B/position of pileup and calls + evidence shares + an
    8 MiB floor.  With a 40 MiB budget a 190 kb request (~24 MiB) fits alone
    but never beside another, so every such request runs in a batch of one;
    with 12 MiB a 150 kb request (~20 MiB) is rejected and a full queue
    answers BUSY, while every accepted request still completes."""
    from pilon_amd.engine import PilonError, Server
    from pilon_amd.synth import serve_requests
    d, ctx, rids = resident
    print("case: admission by memory, rejection, queue limit", flush=True)
    srv = Server(ctx, lanes=2, max_batch=16, mem_budget=40 << 20)
    try:
        for i in range(4):
            srv.submit(i % d.n, 1, 190_000, rids[i % d.n], tag=i)
        res = _wait_all(srv, 4)
        assert all(r["status"] == 0 for r in res)
        assert all(r["batch_size"] == 1 for r in res)
        assert all(20 << 20 < r["est_bytes"] <= 40 << 20 for r in res)
    finally:
        srv.close()
    srv = Server(ctx, lanes=1, max_batch=16, mem_budget=12 << 20, max_queue=4)
    try:
        with pytest.raises(PilonError) as ei:
            srv.submit(0, 1, 150_000, rids[0])
        assert ei.value.status == PL_E_REJECTED
        accepted, busy = 0, 0
        for ci, s, e in serve_requests(d, 2000, 2_000, 20_000, seed=5):
            try:
                srv.submit(ci, s, e, rids[ci])
                accepted += 1
            except PilonError as ex:
                assert ex.status == PL_E_BUSY
                busy += 1
                if busy > 20:
                    break
        assert busy > 0
        res = _wait_all(srv, accepted)
        assert all(r["status"] == 0 for r in res)
        assert all(r["est_bytes"] <= 12 << 20 for r in res)
    finally:
        srv.close()


def test_uploads_refused_while_serving(resident):
    from pilon_amd.engine import PilonError, Server
    d, ctx, rids = resident
    srv = Server(ctx, lanes=1)
    try:
        with pytest.raises(PilonError) as ei:
            ctx.ref_upload(99, "x", "ACGT" * 100)
        assert ei.value.status == PL_E_BUSY
    finally:
        srv.close()
    ctx.ref_upload(99, "x", "ACGT" * 100)   # fine again once the server is closed


def test_close_drains_queue(resident):
    from pilon_amd.engine import Server
    from pilon_amd.synth import serve_requests
    d, ctx, rids = resident
    srv = Server(ctx, lanes=2, max_batch=4)
    for ci, s, e in serve_requests(d, 30, 2_000, 50_000, seed=3):
        srv.submit(ci, s, e, rids[ci])
    srv.close()   # returns after the queued requests ran; no hang, no leak
