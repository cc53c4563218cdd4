"""Local reassembly on the device path (pl_assemble_break: the K-mer graph
built on the GPU, the path search and joins on the host) against the
oracle, exactly: the fix triple and the tandem repeat noted, on gaps,
deletions, partial extensions, heterozygous sites, tandem repeats under
each trSafe / multiClosure setting, no reads, and the edge Pilon throws on."""
import pytest

from pilon_amd.abi import default_config

from . import asm_cases as AC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


def both(engine, c, **over):
    from oracle import asm as A
    opts = dict(c["opts"], **over)
    d = c["draft"]
    bases = d[c["start"] - 1:c["stop"]]
    args = (c["reads"], bases, c.get("contig", d), c["start"], c["stop"], c["bstart"], c["bstop"], c["is_gap"])
    st, sol, loop = A.assemble_break(engine.config, *args, **opts)
    assert st == 0
    return engine.assemble_break(*args, **opts), (sol, loop)


@pytest.mark.parametrize("c", AC.parity_cases(), ids=lambda c: c["name"] + "".join(
    "-%s%s" % (k, v) for k, v in sorted(c["opts"].items())))
def test_assemble_break_matches_oracle(engine, c):
    print("case: assemble_break", c["name"], c["opts"], flush=True)
    got, ref = both(engine, c)
    assert got == ref


def test_snp_fixed_bases_differ_from_the_contig(engine):
    """trimPatch compares the patch with the region's current bases and
    with the contig as read: a base fixed earlier in the region (here the
    first base after the gap) matches either way."""
    print("case: assemble_break with a fixed base", flush=True)
    for which in ("bases", "both"):
        c = AC.fixed_base_case(which)
        got, ref = both(engine, c)
        assert got == ref


def test_gap_at_region_start_raises_edge(engine):
    from pilon_amd.engine import PilonError
    print("case: assemble_break edge", flush=True)
    c = AC.gap_case(1)
    d = c["draft"]
    with pytest.raises(PilonError) as e:
        engine.assemble_break(c["reads"], d, d, 1, len(d), 1, 100, True)
    assert e.value.status == 4


def test_bad_arguments_are_refused(engine):
    from pilon_amd.engine import PilonError
    print("case: assemble_break bad arguments", flush=True)
    c = AC.gap_case(1)
    d = c["draft"]
    with pytest.raises(PilonError):
        engine.assemble_break(c["reads"], d, d, 1, len(d), 900, 800, True)      # break stop before start
    with pytest.raises(PilonError):
        engine.assemble_break(c["reads"], d, d, 1, len(d), 801, 900, True, k=1000)
    with pytest.raises(ValueError):
        engine.assemble_break(c["reads"], d[:10], d, 1, len(d), 801, 900, True)
