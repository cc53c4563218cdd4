"""The oracle's Assembler k-mer graph (addReads + buildGraph) against a
literal Python restatement with a dict of per-K-mer pileups (the HashMap of
Assembler.scala), on hand cases, random reads with IUPAC codes and absent
qualities, and synthetic reads at K = 47."""
import numpy as np
import pytest

from pilon_amd.abi import PL_RF_EXOTIC, PL_RF_NO_QUAL, default_config
from pilon_amd.reads import pack_reads, unpack_2bit, unpack_nib

NIB = "=ACMGRSVTWYHKDBN"
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def oracle():
    from oracle import oracle as O
    return O


def read_strings(b, default_qual):
    exo = {int(r): int(o) for r, o in zip(b.exo_read, b.exo_off)}
    for i in range(b.n):
        s0, s1 = int(b.seq_off[i]), int(b.seq_off[i + 1])
        L = s1 - s0
        if i in exo:
            seq = "".join(NIB[c] for c in unpack_nib(b.nib, exo[i], L))
        else:
            seq = "".join("ACGT"[c] for c in unpack_2bit(b.bases2, s0, L))
        q = [default_qual] * L if b.flag[i] & PL_RF_NO_QUAL else [int(x) for x in b.quals[s0:s1]]
        yield seq, q, int(b.mapq[i])


def dict_graph(cfg, b, K, min_depth):
    piles = {}
    for seq, q, mq in read_strings(b, cfg.default_qual):
        L = len(seq)
        if L <= K:
            continue
        rc = "".join(COMP.get(c, c) for c in reversed(seq))
        for s, qq in ((seq, q), (rc, q[::-1])):
            for o in range(L - K):
                p = piles.setdefault(s[o:o + K], [[0] * 4, [0] * 4, [0]])
                c, qual = s[o + K], qq[o + K]
                if c in "ACGT" and qual >= cfg.min_qual:       # PileUp.add
                    bi = "ACGT".index(c)
                    p[0][bi] += 1
                    p[1][bi] += qual * (mq + 1)
                    p[2][0] += qual
    out = []
    for k in sorted(piles, key=lambda x: x.encode()):
        bc, qs, qsum = piles[k]
        if sum(bc) < min_depth:
            continue
        src = qs if qsum[0] > 0 else bc                              # BaseCall: qualSum when qSum > 0
        order = sorted(range(4), key=lambda i: -src[i])             # stable: ties keep A<C<G<T
        b0, b1 = order[0], order[1]
        total = sum(qs)
        homo_s = qs[b0] - (total - qs[b0])
        half = total // 2
        het_s = total - abs(half - qs[b0]) - abs(half - qs[b1])
        homo = homo_s >= het_s
        out.append((k, "ACGT"[b0], None if homo else "ACGT"[b1]))
    return out


def test_hand_case():
    """Five copies of ACGTTGCA (K = 4): forward k-mer ACGT is followed by T,
    GTTG by C; on the reverse strand (TGCAACGT) TGCA is followed by A."""
    O = oracle()
    reads = [dict(pos=1, cigar="8M", seq="ACGTTGCA", qual=[30] * 8, flag=0, mapq=20) for _ in range(5)]
    st, g = O.kmer_graph(default_config(), pack_reads(reads), K=4, min_depth=5)
    assert st == 0
    d = {k: (n, a) for k, n, a in g}
    assert d["ACGT"] == ("T", None) and d["GTTG"] == ("C", None) and d["TGCA"] == ("A", None)
    assert len(g) == 8    # 4 forward + 4 reverse k-mers, each seen 5 times


def test_heterozygous_next_base():
    """A k-mer followed by C in 5 reads and G in 4 is not homozygous: kGraph C, altGraph G."""
    O = oracle()
    reads = [dict(pos=1, cigar="6M", seq="AAAAAC", qual=[30] * 6, mapq=60) for _ in range(5)]
    reads += [dict(pos=1, cigar="6M", seq="AAAAAG", qual=[30] * 6, mapq=60) for _ in range(4)]
    st, g = O.kmer_graph(default_config(), pack_reads(reads), K=5, min_depth=5)
    d = {k: (n, a) for k, n, a in g}
    assert d["AAAAA"] == ("C", "G")
    assert st == 0


@pytest.mark.parametrize("K", [3, 7, 15, 47])
def test_random_reads_against_dict(K):
    rng = np.random.default_rng(K)
    cfg = default_config()
    reads = []
    for i in range(300):
        L = int(rng.integers(max(1, K - 3), K + 40))
        alpha = "ACGT" if rng.random() < 0.7 else "ACGTNRY"
        seq = "".join(rng.choice(list(alpha), L))
        qual = None if rng.random() < 0.1 else [int(x) for x in rng.integers(0, 41, L)]
        reads.append(dict(pos=int(rng.integers(1, 500)), cigar=f"{L}M", seq=seq, qual=qual,
                          mapq=int(rng.integers(0, 61))))
    # a few copies make k-mers reach depth 5
    reads += reads[:40] * 4
    b = pack_reads(reads)
    st, g = oracle().kmer_graph(cfg, b, K=K, min_depth=5)
    assert st == 0
    assert g == dict_graph(cfg, b, K, 5)
    assert g


def test_synthetic_reads_k47():
    from pilon_amd.synth import SynthData, preset
    cfg = default_config()
    p = preset("c2", scale=0.004)
    p.update({"n_rate": 0.002})
    d = SynthData(p)
    b = d.batch(0).slice(0, 3000)
    st, g = oracle().kmer_graph(cfg, b, K=47, min_depth=5)
    assert st == 0 and len(g) > 1000
    assert g == dict_graph(cfg, b, 47, 5)
    assert any(a is not None for _, _, a in g)


def test_limits():
    O = oracle()
    b = pack_reads([dict(pos=1, cigar="10M", seq="ACGTACGTAC", qual=[30] * 10)])
    assert O.kmer_graph(default_config(), b, K=49)[0] == 1
    assert O.kmer_graph(default_config(), b, K=0)[0] == 1
