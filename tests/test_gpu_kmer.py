"""The assembler's K-mer graph on the device (pl_kmer_graph) against the
oracle's, exactly: hand cases, random reads with IUPAC codes and absent
qualities at several K, reads of a synthetic contig (K = 47, the
assembler's default), and the refusal past the 100,000-read prune point."""
import numpy as np
import pytest

from pilon_amd.abi import default_config
from pilon_amd.reads import pack_reads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine():
    from pilon_amd.engine import Context
    ctx = Context(default_config(), device=0)
    yield ctx
    ctx.close()


def test_kmer_hand_cases(engine):
    from oracle import oracle as O
    print("case: kmer hand cases", flush=True)
    reads = [dict(pos=1, cigar="8M", seq="ACGTTGCA", qual=[30] * 8, mapq=20) for _ in range(5)]
    reads += [dict(pos=1, cigar="6M", seq="AAAAAC", qual=[30] * 6, mapq=60) for _ in range(5)]
    reads += [dict(pos=1, cigar="6M", seq="AAAAAG", qual=[30] * 6, mapq=60) for _ in range(4)]
    b = pack_reads(reads)
    for K in (4, 5):
        got, _ = engine.kmer_graph(b, K=K, min_depth=5)
        st, ref = O.kmer_graph(engine.config, b, K=K, min_depth=5)
        assert st == 0 and got == ref and got


@pytest.mark.parametrize("K", [3, 7, 15, 31, 47, 48])
def test_kmer_random_reads(engine, K):
    from oracle import oracle as O
    print(f"case: kmer random reads K={K}", flush=True)
    rng = np.random.default_rng(100 + K)
    reads = []
    for _ in range(400):
        L = int(rng.integers(max(1, K - 3), K + 60))
        alpha = "ACGT" if rng.random() < 0.7 else "ACGTNRYM"
        seq = "".join(rng.choice(list(alpha), L))
        qual = None if rng.random() < 0.1 else [int(x) for x in rng.integers(0, 41, L)]
        reads.append(dict(pos=int(rng.integers(1, 500)), cigar=f"{L}M", seq=seq, qual=qual,
                          mapq=int(rng.integers(0, 61))))
    reads += reads[:60] * 4
    b = pack_reads(reads)
    got, _ = engine.kmer_graph(b, K=K, min_depth=5)
    st, ref = O.kmer_graph(engine.config, b, K=K, min_depth=5)
    assert st == 0
    assert got == ref and got


@pytest.mark.parametrize("n_reads,min_depth", [(20000, 5), (99999, 3)])
def test_kmer_synthetic_reads(engine, n_reads, min_depth):
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    print(f"case: kmer synthetic reads n={n_reads} K=47", flush=True)
    p = preset("c2", scale=0.04)
    p.update({"n_rate": 0.002})
    d = SynthData(p)
    b = d.batch(0).slice(0, n_reads)
    got, ms = engine.kmer_graph(b, K=47, min_depth=min_depth)
    st, ref = O.kmer_graph(engine.config, b, K=47, min_depth=min_depth)
    assert st == 0 and len(ref) > 10000
    assert got == ref
    assert ms > 0


@pytest.mark.parametrize("K", [31, 47])
def test_kmer_word_order_path(engine, K, monkeypatch):
    """The order by the K-mer words themselves (taken when two K-mers share
    a hash; PILON_HIP_KMER_WORD_SORT takes it without one) gives the same
    graph as the order by hash, and the oracle's."""
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    print(f"case: kmer word order K={K}", flush=True)
    p = preset("c2", scale=0.04)
    p.update({"n_rate": 0.002})
    b = SynthData(p).batch(0).slice(0, 20000)
    by_hash, _ = engine.kmer_graph(b, K=K, min_depth=3)
    monkeypatch.setenv("PILON_HIP_KMER_WORD_SORT", "1")
    by_words, _ = engine.kmer_graph(b, K=K, min_depth=3)
    monkeypatch.delenv("PILON_HIP_KMER_WORD_SORT")
    st, ref = O.kmer_graph(engine.config, b, K=K, min_depth=3)
    assert st == 0 and len(ref) > 1000
    assert by_words == by_hash == ref


def test_kmer_prune_point_refused(engine):
    from pilon_amd.engine import PilonError
    from pilon_amd.synth import SynthData, preset
    print("case: kmer 100000 reads refused", flush=True)
    d = SynthData(preset("c2", scale=0.04))
    b = d.batch(0).slice(0, 100000)
    with pytest.raises(PilonError) as e:
        engine.kmer_graph(b, K=47)
    assert e.value.status == 1
