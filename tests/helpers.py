"""Shared comparison helpers for the parity tests."""
import numpy as np

from pilon_amd.abi import FIELDS


def assert_same_region(got_pile, got_calls, ref, what=""):
    """Bit-exact equality of every PileUp field and every pl_call."""
    assert got_pile.shape == ref["pile"].shape, (what, got_pile.shape, ref["pile"].shape)
    for f in range(len(FIELDS)):
        if not np.array_equal(got_pile[f], ref["pile"][f]):
            bad = np.nonzero(got_pile[f] != ref["pile"][f])[0]
            raise AssertionError(f"{what}: field {FIELDS[f]} differs at {len(bad)} positions, first {bad[:5]}: "
                                 f"got {got_pile[f][bad[:5]]} want {ref['pile'][f][bad[:5]]}")
    a = got_calls.view(np.uint64)
    b = ref["calls"].view(np.uint64)
    if not np.array_equal(a, b):
        bad = np.nonzero(a != b)[0]
        raise AssertionError(f"{what}: calls differ at {len(bad)} positions, first {bad[:5]}: "
                             f"got {got_calls[bad[:3]]} want {ref['calls'][bad[:3]]}")


STAT_KEYS = ["read_count", "base_count", "coverage", "min_depth", "n_confirmed", "snps", "amb", "ins",
             "ins_bases", "dels", "del_bases", "n_unknown_ops", "aligned_bases"]


def assert_same_stats(got, ref, what=""):
    diff = {k: (got[k], ref[k]) for k in STAT_KEYS if got[k] != ref[k]}
    assert not diff, f"{what}: stats differ {diff}"
