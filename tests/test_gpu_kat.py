"""The hand-derived known-answer cases through libpilonhip.so on the GPU."""
import pytest

from pilon_amd.abi import default_config
from tests import kat_cases as K
from tests.helpers import assert_same_region, assert_same_stats
from tests.test_oracle_kat import check_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


@pytest.mark.parametrize("make", K.ALL, ids=[f.__name__ for f in K.ALL])
def test_kat_gpu(engine, make):
    from pilon_amd.engine import PileUpRegion
    from oracle import oracle as O
    case = make()
    print(f"case: {make.__name__}", flush=True)
    cfg = default_config(**case.get("cfg", {}))
    own = None
    if case.get("cfg"):
        from pilon_amd.engine import Context
        engine = own = Context(cfg, device=0)
    try:
        engine.ref_upload(100, "ctg", case["ref"].encode())
        pr = PileUpRegion(engine, 100, case["start"], case["stop"])
        pr.addReads(K.batch_of(case))
        pr.postProcess()
        changes, bases = pr.fix()
        wide = engine.wide_columns()
        pr.close()
    finally:
        if own is not None:
            own.close()
    res = {"status": 0, "pile": pr.pile, "calls": pr.calls, "stats": pr.stats, "changes": changes, "bases": bases}
    check_case(res, case)
    ref = O.run_region(cfg, case["ref"], case["start"], case["stop"], [K.batch_of(case)], name="ctg")
    assert_same_region(pr.pile, pr.calls, ref, make.__name__)
    assert_same_stats(pr.stats, ref["stats"])
    # exact Long qualSums of the columns beyond int32 (the planes hold their low bits)
    want = {i: tuple(int(x) for x in ref["bcq"][4:8, i]) for i in range(ref["size"])
            if any(not -2 ** 31 <= int(x) < 2 ** 31 for x in ref["bcq"][4:8, i])}
    assert wide == want, make.__name__
    assert (case.get("wide") is not None) == bool(want)


def test_edge_gpu(engine):
    from pilon_amd.engine import PileUpRegion, PilonError
    print("case: insertion slides before region start", flush=True)
    case = K.case_insertion_rotation(start=14)
    engine.ref_upload(101, "ctg", case["ref"].encode())
    pr = PileUpRegion(engine, 101, 14, 40)
    pr.addReads(K.batch_of(case))
    with pytest.raises(PilonError) as e:
        pr.postProcess()
    assert e.value.status == 4


@pytest.mark.parametrize("make", K.LONG, ids=[f.__name__ for f in K.LONG])
def test_kat_long_read_gpu(make):
    """The long-read cases through pl_region_add_reads_typed, one call per BAM."""
    from pilon_amd.abi import BAM_TYPES
    from pilon_amd.engine import Context, PileUpRegion
    from oracle import oracle as O
    case = make()
    print(f"case: {make.__name__}", flush=True)
    cfg = default_config(**case.get("cfg", {}))
    sets = K.sets_of(case)
    ctx = Context(cfg, device=0)
    try:
        ctx.ref_upload(100, "ctg", case["ref"].encode())
        pr = PileUpRegion(ctx, 100, case["start"], case["stop"])
        for b, t in sets:
            pr.addReads(b, bam_type=t)
        pr.postProcess()
        changes, bases = pr.fix()
        pr.close()
    finally:
        ctx.close()
    res = {"status": 0, "pile": pr.pile, "calls": pr.calls, "stats": pr.stats, "changes": changes, "bases": bases}
    check_case(res, case)
    ref = O.run_region(cfg, case["ref"], case["start"], case["stop"], [b for b, _ in sets], name="ctg",
                       set_types=[BAM_TYPES[t] for _, t in sets])
    assert_same_region(pr.pile, pr.calls, ref, make.__name__)
    assert_same_stats(pr.stats, ref["stats"])
    assert changes == ref["changes"] and bases == ref["bases"]
