"""The kernel core built for the host (same window decomposition, runs,
deferral and Pass-1 code as the gfx950 kernels) against the oracle."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from tests import corehost as CH
from tests import kat_cases as K
from tests.helpers import assert_same_region, assert_same_stats


def oracle():
    from oracle import oracle as O
    return O


@pytest.mark.parametrize("make", K.ALL, ids=[f.__name__ for f in K.ALL])
@pytest.mark.parametrize("W", [64, 1024])
def test_kat_core_host(make, W):
    case = make()
    b = K.batch_of(case)
    cfg = default_config(**case.get("cfg", {}))
    ref = oracle().run_region(cfg, case["ref"], case["start"], case["stop"], [b], name="ctg")
    got = CH.run(cfg, case["ref"].encode(), b, [(case["start"], case["stop"])], W=W, name="ctg")[0]
    assert got["status"] == 0
    assert_same_region(got["pile"], got["calls"], ref, make.__name__)
    assert_same_stats(got["stats"], ref["stats"])
    assert got["changes"] == ref["changes"] and got["bases"] == ref["bases"]


@pytest.mark.parametrize("make", K.LONG, ids=[f.__name__ for f in K.LONG])
def test_kat_long_read_core_host(make):
    """The long-read cases: the reads walked with their longReadType, and
    excludeMotifs on the change list."""
    from pilon_amd.abi import BAM_TYPES
    case = make()
    cfg = default_config(**case.get("cfg", {}))
    sets = K.sets_of(case)
    ref = oracle().run_region(cfg, case["ref"], case["start"], case["stop"], [b for b, _ in sets], name="ctg",
                              set_types=[BAM_TYPES[t] for _, t in sets])
    b, lrt, kinds = K.merged_of(case)
    got = CH.run(cfg, case["ref"].encode(), b, [(case["start"], case["stop"])], W=64, name="ctg", lrt=lrt,
                 set_kinds=kinds)[0]
    assert got["status"] == 0
    assert_same_region(got["pile"], got["calls"], ref, make.__name__)
    assert_same_stats(got["stats"], ref["stats"])
    assert got["changes"] == ref["changes"] and got["bases"] == ref["bases"]


@pytest.mark.parametrize("lr", ["nanopore", "pacbio"])
def test_synthetic_long_read_core_host(lr):
    """A c2-shaped contig (error indels, N runs) piled as one long-read BAM:
    every homopolymer and CC_GG motif of the random draft meets the reads'
    indels and bases."""
    from pilon_amd.abi import BAM_TYPES
    from pilon_amd.synth import SynthData, preset
    p = preset("c2", scale=0.04)
    p.update({"n_rate": 0.001, "err_indel_rate": 0.004})
    d = SynthData(p)
    cfg = default_config(longread=1)
    draft, b = d.draft(0), d.batch(0)
    L = len(draft)
    regions = [(1, L // 2), (L // 2 + 1, L), (7, L - 5)]
    lrt = np.full(b.n, K.LR_TYPE[lr], np.uint8)
    got = CH.run(cfg, draft, b, regions, W=512, name=d.names[0], lrt=lrt, set_kinds=2 if lr == "nanopore" else 0)
    for (s, e), g in zip(regions, got):
        ref = oracle().run_region(cfg, draft, s, e, [b], name=d.names[0], set_types=[BAM_TYPES[lr]])
        assert g["status"] == ref["status"] == 0
        assert_same_region(g["pile"], g["calls"], ref, f"{lr} {s}-{e}")
        assert_same_stats(g["stats"], ref["stats"])
        assert g["changes"] == ref["changes"] and g["bases"] == ref["bases"]
    # the long-read rules changed the pileup: indel quality sums differ from a frags run
    plain = oracle().run_region(cfg, draft, 1, L // 2, [b], name=d.names[0])
    assert not np.array_equal(plain["pile"], got[0]["pile"])


@pytest.mark.parametrize("name,scale,extra", [
    ("c1", 1.0, {}),
    ("c2", 0.04, {"n_rate": 0.001, "err_indel_rate": 0.002}),
    ("c5", 0.004, {}),
])
def test_synthetic_core_host(name, scale, extra):
    p = __import__("pilon_amd.synth", fromlist=["preset"]).preset(name, scale=scale)
    p.update(extra)
    from pilon_amd.synth import SynthData
    d = SynthData(p)
    cfg = default_config()
    draft, b = d.draft(0), d.batch(0)
    L = len(draft)
    regions = [(1, L // 3), (L // 3 + 1, 2 * L // 3), (2 * L // 3 + 1, L), (1, L)]
    for W in (256, 2048):
        got = CH.run(cfg, draft, b, regions, W=W, name=d.names[0])
        for (s, e), g in zip(regions, got):
            ref = oracle().run_region(cfg, draft, s, e, [b], name=d.names[0])
            assert g["status"] == ref["status"] == 0
            assert_same_region(g["pile"], g["calls"], ref, f"{name} {s}-{e} W={W}")
            assert_same_stats(g["stats"], ref["stats"])
            assert g["changes"] == ref["changes"] and g["bases"] == ref["bases"]


def test_edge_detected_by_core_host():
    case = K.case_insertion_rotation(start=14)
    got = CH.run(default_config(), case["ref"].encode(), K.batch_of(case), [(14, 40)], W=64)[0]
    assert got["status"] == 4
