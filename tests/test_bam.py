"""The BAM reader (libpilonbam.so) against records encoded literally from
the SAMv1 specification (tests/bamspec.py): the decode rules the pileup path
depends on, BGZF framing and error cases, and round trips of synthetic
workloads that must give the oracle the same regions as the original
batches."""
import struct

import numpy as np
import pytest

from pilon_amd.abi import PL_RF_EXOTIC, PL_RF_MATE_SAME_REF, PL_RF_NO_QUAL, default_config
from pilon_amd.bam import BamError, BamFile
from pilon_amd.reads import unpack_2bit
from tests import bamspec as S


def w(n, op):
    return (n << 4) | "MIDNSHP=X".index(op)


def known_records():
    cg = b"CGBI" + struct.pack("<I", 3) + struct.pack("<3I", w(4, "M"), w(2, "D"), w(6, "M"))
    recs = [
        # ref 0: a proper pair mate (mate on the same reference)
        S.record(0, 9, [w(5, "M")], "ACGTA", [30, 31, 32, 33, 34], flag=0x63, next_ref=0, tlen=200, name="a"),
        # non-ACGT codes, QUAL absent, mate on another reference, a flag bit above 0x800
        S.record(0, 9, [w(3, "M"), w(1, "I"), w(2, "M")], "ACNGTM", None, flag=0x1891, next_ref=1, tlen=-50,
                 mapq=7, name="b"),
        # placed but unmapped, SEQ "*"
        S.record(0, 20, [], "", None, flag=0x4, next_ref=-1, name="c"),
        # a CIGAR carried by the CG:B,I tag (kSmN placeholder in the record)
        S.record(0, 30, [w(10, "S"), w(10, "N")], "ACGTACGTAC", [20] * 10, flag=0, name="d",
                 tags=b"NMi" + struct.pack("<i", 1) + cg + b"XZZhello\0"),
        # ref 1
        S.record(1, 0, [w(2, "S"), w(3, "M")], "GGACG", [40] * 5, flag=0x10, name="e"),
        # unplaced: never handed over
        S.record(-1, -1, [], "ACGT", [10] * 4, flag=0x4, name="f"),
    ]
    return S.header([("chrA", 1000), ("chrB", 500)]) + b"".join(recs)


def test_eof_block_is_the_spec_marker():
    assert S.bgzf_block(b"") == S.SPEC_EOF


@pytest.mark.parametrize("block,threads", [(0xFF00, 1), (7, 1), (7, 4), (1000, 3)])
def test_known_records(tmp_path, block, threads):
    p = tmp_path / "k.bam"
    p.write_bytes(S.bgzf(known_records(), block=block))
    with BamFile(p, threads=threads) as b:
        assert b.names == ["chrA", "chrB"] and b.lens == [1000, 500]
        st = b.stats
        assert st["records"] == 6 and st["placed"] == 5 and st["exotic"] == 1 and st["eof_marker"] == 1
        a = b.batch(0)
        assert a.pos.tolist() == [10, 10, 21, 31]                      # getAlignmentStart = pos + 1
        assert a.tlen.tolist() == [200, -50, 0, 0]
        assert a.mapq.tolist() == [60, 7, 60, 60]
        assert a.flag.tolist() == [0x63 | PL_RF_MATE_SAME_REF, 0x891 | PL_RF_NO_QUAL | PL_RF_EXOTIC,
                                   0x4 | PL_RF_NO_QUAL, 0]
        assert a.cigar_off.tolist() == [0, 1, 4, 4, 7]
        assert a.cigar.tolist() == [w(5, "M"), w(3, "M"), w(1, "I"), w(2, "M"), w(4, "M"), w(2, "D"), w(6, "M")]
        assert a.seq_off.tolist() == [0, 5, 11, 11, 21]
        assert unpack_2bit(a.bases2, 0, 5).tolist() == [0, 1, 2, 3, 0]
        assert unpack_2bit(a.bases2, 11, 10).tolist() == [0, 1, 2, 3, 0, 1, 2, 3, 0, 1]
        assert a.quals[0:5].tolist() == [30, 31, 32, 33, 34]
        assert a.quals[5:11].tolist() == [-1] * 6                      # 0xFF as a Java byte
        assert a.exo_read.tolist() == [1] and a.exo_off.tolist() == [0]
        # "ACNGTM": A=1 C=2 N=15 G=4 T=8 M=3, first base in the high nibble
        assert a.nib[:3].tolist() == [0x12, 0xF4, 0x83]
        e = b.batch(1)
        assert e.pos.tolist() == [1] and e.flag.tolist() == [0x10]
        assert unpack_2bit(e.bases2, 0, 5).tolist() == [2, 2, 0, 1, 2]


def test_header_and_records_split_over_many_blocks(tmp_path):
    text = b"@HD\tVN:1.6\tSO:coordinate\n" + b"@CO\t" + b"x" * 200000 + b"\n"
    data = S.header([("c", 100)], text=text) + S.record(0, 4, [w(4, "M")], "TTGA", [9, 9, 9, 9])
    p = tmp_path / "h.bam"
    p.write_bytes(S.bgzf(data, block=5000))
    with BamFile(p, threads=2) as b:
        a = b.batch(0)
        assert a.pos.tolist() == [5] and unpack_2bit(a.bases2, 0, 4).tolist() == [3, 3, 2, 0]


def test_no_eof_marker_is_reported(tmp_path):
    p = tmp_path / "n.bam"
    p.write_bytes(S.bgzf(known_records(), eof=False))
    with BamFile(p) as b:
        assert b.stats["eof_marker"] == 0 and b.batch(0).n == 4


def test_errors(tmp_path):
    good = S.bgzf(known_records(), block=300)
    cases = {}
    bad = bytearray(good)
    bad[40] ^= 0x55                                                    # inside the first block's deflate data
    cases["crc"] = bytes(bad)
    cases["truncated"] = good[:-40]
    cases["not bgzf"] = b"\x1f\x8b\x08\x00" + b"\0" * 40
    cases["not bam"] = S.bgzf(b"SAM\x01" + b"\0" * 64)
    unsorted = S.header([("c", 100)]) + S.record(0, 50, [w(2, "M")], "AC", [9, 9]) + \
        S.record(0, 10, [w(2, "M")], "AC", [9, 9])
    cases["unsorted"] = S.bgzf(unsorted)
    h = S.header([("c", 100)])
    short = bytearray(h + S.record(0, 5, [w(8, "M")], "ACGTACGT", [9] * 8))
    struct.pack_into("<I", short, len(h), 40)                          # block_size below its contents
    cases["block_size"] = S.bgzf(bytes(short))
    for name, blob in cases.items():
        p = tmp_path / f"{name.replace(' ', '_')}.bam"
        p.write_bytes(blob)
        with pytest.raises(BamError) as e:
            BamFile(p)
        assert e.value.status == 1, name


def _synth_roundtrip(tmp_path, name, scale, extra, threads, monkeypatch, slice_kib):
    from oracle import oracle as O
    from pilon_amd.synth import SynthData, preset
    from tests.helpers import assert_same_region
    p = preset(name, scale=scale)
    p.update(extra)
    d = SynthData(p)
    refs = [(d.names[i], len(d.draft(i))) for i in range(d.n)]
    batches = {i: d.batch(i) for i in range(d.n)}
    path = tmp_path / f"{name}.bam"
    nbytes = S.write_bam(path, refs, batches)
    cfg = default_config()
    monkeypatch.setenv("PILON_BAM_SLICE_KIB", str(slice_kib))
    with BamFile(path, threads=threads) as b:
        assert b.names == [r[0] for r in refs]
        for i in range(d.n):
            S.same_batch(batches[i], b.batch(i), f"{name} ref {i}")
        # the oracle sees the same regions through the decoded batch
        i = 0
        L = len(d.draft(i))
        ref = O.run_region(cfg, d.draft(i), 1, L, [batches[i]], name=d.names[i])
        got = O.run_region(cfg, d.draft(i), 1, L, [b.reads(i)], name=d.names[i])
        assert_same_region(got["pile"], got["calls"], ref, f"{name} via BAM")
        assert got["changes"] == ref["changes"] and got["bases"] == ref["bases"]
        return b.stats, nbytes


def test_synthetic_roundtrip_c2(tmp_path, monkeypatch):
    # slices of 256 KiB: records cut at slice ends are carried to the next
    st, nbytes = _synth_roundtrip(tmp_path, "c2", 0.02, {"n_rate": 0.001}, 3, monkeypatch, 256)
    assert st["exotic"] > 0 and st["placed"] > 10000
    assert st["bam_bytes"] == nbytes and nbytes > 40 * (256 << 10)


def test_synthetic_roundtrip_many_contigs(tmp_path, monkeypatch):
    st, nbytes = _synth_roundtrip(tmp_path, "c3", 0.002, {}, 4, monkeypatch, 1024)
    assert st["bam_bytes"] == nbytes


def test_synthetic_roundtrip_segmented_delimiting(tmp_path, monkeypatch):
    """Slices of 8 MiB on five threads: each slice's records are delimited
    per segment from a guessed start, checked against the previous
    segment's end."""
    st, nbytes = _synth_roundtrip(tmp_path, "c2", 0.02, {"n_rate": 0.001}, 5, monkeypatch, 8192)
    assert st["bam_bytes"] == nbytes and nbytes > 2 * (8192 << 10)


def test_lookalike_record_chains_in_quals(tmp_path, monkeypatch):
    """QUAL bytes that hold three well-formed records in a row: a segment
    whose start falls before them guesses them as its first record, the
    check against the previous segment's end rejects the guess, and the
    chain from the true start gives the same batch as one thread does."""
    fake = b"".join(S.record(0, 7 + k, [w(1, "M")], "", [], name="x") for k in range(3))
    fake_q = list(fake)
    n_rec, L = 60000, 4 * len(fake_q)
    seq = "ACGT" * (L // 4)
    q = (fake_q * 4)[:L]
    body = S.header([("c", 10 ** 7)]) + b"".join(S.record(0, 10 * k, [w(L, "M")], seq, q, name="r%d" % k)
                                                 for k in range(n_rec))
    p = tmp_path / "lookalike.bam"
    p.write_bytes(S.bgzf(body, level=1))
    assert len(body) > (16 << 20)
    monkeypatch.setenv("PILON_BAM_SLICE_KIB", "16384")
    with BamFile(p, threads=7) as b:
        many = b.batch(0)
        assert b.stats["placed"] == n_rec
    with BamFile(p, threads=1) as b:
        one = b.batch(0)
    S.same_batch(one, many, "lookalike chains")
    assert list(many.pos[:3]) == [1, 11, 21] and bytes(many.quals[:len(fake)].astype("uint8")) == fake


@pytest.mark.parametrize("cap_kib", [128, 8192])
def test_stream_hands_over_every_record(tmp_path, monkeypatch, cap_kib):
    """pl_bam_stream_*: the header as pl_bam_open parses it, then every record
    in file order, whole, in slices of the caller's size (a record cut at a
    slice's end comes first in the next one): their refIDs and positions
    are the decoded batches'."""
    import struct
    from pilon_amd.bam import BamStream
    from pilon_amd.synth import SynthData, preset
    d = SynthData(preset("c2", scale=0.01))
    refs = [(d.names[i], len(d.draft(i))) for i in range(d.n)]
    batches = {i: d.batch(i) for i in range(d.n)}
    path = tmp_path / "s.bam"
    S.write_bam(path, refs, batches)
    with BamStream(path, threads=3) as s:
        assert s.names == [r[0] for r in refs] and s.lens == [r[1] for r in refs]
        got = []
        for buf, offs in s.slices(cap=cap_kib << 10):
            raw = buf.tobytes()
            for o in offs:
                bs, ref, pos = struct.unpack_from("<Iii", raw, int(o))
                assert int(o) + 4 + bs <= len(raw)
                got.append((ref, pos + 1))
        st = s.stats
        total = s.total_bytes
    want = [(i, int(p)) for i in range(d.n) for p in batches[i].pos]
    assert got == want and st["records"] == len(want) and st["eof_marker"] == 1
    assert st["bam_bytes"] == total
    with pytest.raises(BamError):
        with BamStream(path) as s:
            s.soa(0)


def test_native_writer_matches_the_spec_writer(tmp_path):
    """The generator's BAM writer (used for workload-size files) produces the
    same bytes as the spec-literal writer at the same level and block size."""
    from pilon_amd.synth import SynthData, preset
    p = preset("c2", scale=0.004)
    p.update({"n_rate": 0.002})
    d = SynthData(p)
    a, b = tmp_path / "native.bam", tmp_path / "spec.bam"
    n1 = d.write_bam(a, level=1)
    n2 = S.write_bam(b, [(d.names[i], len(d.draft(i))) for i in range(d.n)], {i: d.batch(i) for i in range(d.n)},
                     level=1)
    assert n1 == n2
    assert a.read_bytes() == b.read_bytes()


# ------------------------------------------------------------ BAI and queries
def overlap_filter(batch, start, stop):
    """Reads of a batch whose span [pos, pos + reference length - 1] (one base
    for unmapped or zero-length alignments) overlaps [start, stop] (1-based;
    start <= 0 from the first base): htsjdk's queryOverlapping."""
    keep = []
    for i in range(batch.n):
        cig = batch.cigar[int(batch.cigar_off[i]):int(batch.cigar_off[i + 1])]
        rl = sum(int(x) >> 4 for x in cig if int(x) & 15 in (0, 2, 3, 7, 8))
        ps = int(batch.pos[i])
        pe = ps if (int(batch.flag[i]) & 4 or rl == 0) else ps + rl - 1
        if ps <= max(stop, 1) and pe >= max(start, 1):
            keep.append(i)
    return keep


def same_reads(q, whole, idx, what):
    """q holds exactly reads idx of whole, in order."""
    from pilon_amd.reads import ReadBatch
    assert q.n == len(idx), (what, q.n, len(idx))
    if not idx:
        return
    sub = whole.slice_rows(idx) if hasattr(whole, "slice_rows") else None
    for i_q, i_w in enumerate(idx):
        for k in ("pos", "tlen", "flag", "mapq"):
            assert getattr(q, k)[i_q] == getattr(whole, k)[i_w], (what, k, i_q)
        cq = q.cigar[int(q.cigar_off[i_q]):int(q.cigar_off[i_q + 1])]
        cw = whole.cigar[int(whole.cigar_off[i_w]):int(whole.cigar_off[i_w + 1])]
        assert np.array_equal(cq, cw), (what, "cigar", i_q)
        sq0, sq1 = int(q.seq_off[i_q]), int(q.seq_off[i_q + 1])
        sw0, sw1 = int(whole.seq_off[i_w]), int(whole.seq_off[i_w + 1])
        assert sq1 - sq0 == sw1 - sw0
        if not int(whole.flag[i_w]) & PL_RF_NO_QUAL:
            assert np.array_equal(q.quals[sq0:sq1], whole.quals[sw0:sw1]), (what, "quals", i_q)
        if not int(whole.flag[i_w]) & PL_RF_EXOTIC:
            assert np.array_equal(unpack_2bit(q.bases2, sq0, sq1 - sq0), unpack_2bit(whole.bases2, sw0, sw1 - sw0))
    assert len(q.exo_read) == sum(1 for i in idx if int(whole.flag[i]) & PL_RF_EXOTIC), what
    del sub


def spanning_batches():
    """Two references; reads spread over 200 kb with a 20 kb deletion read
    (bin far above its neighbours), unmapped-placed and zero-length reads,
    non-ACGT codes, absent QUAL."""
    from pilon_amd.reads import pack_reads
    rng = np.random.default_rng(21)
    reads = []
    for k in range(1500):
        p = int(rng.integers(1, 200_000))
        seq = "".join(rng.choice(list("ACGT"), 60))
        reads.append(dict(pos=p, cigar="60M", seq=seq, qual=[int(x) for x in rng.integers(2, 41, 60)], flag=0x1 | 0x2,
                          mapq=60, tlen=300))
    reads.append(dict(pos=1000, cigar="10M20000D10M", seq="A" * 20, qual=[30] * 20, flag=0, mapq=50, tlen=0))
    reads.append(dict(pos=150_000, cigar="*", seq="ACGT", qual=None, flag=0x4, mapq=0, tlen=0))
    reads.append(dict(pos=90_000, cigar="5S", seq="ACGTN", qual=[20] * 5, flag=0, mapq=0, tlen=0))
    reads.append(dict(pos=120_000, cigar="30M", seq="ACGTNACGTRACGTACGTACGTACGTACGT", qual=None, flag=0, mapq=9,
                      tlen=0))
    b0 = pack_reads(reads)
    b1 = pack_reads([dict(pos=int(p), cigar="40M", seq="C" * 40, qual=[25] * 40, flag=0, mapq=60, tlen=0)
                     for p in sorted(rng.integers(1, 50_000, 300))])
    return [("chrA", 250_000), ("chrB", 60_000)], {0: b0, 1: b1}


@pytest.mark.parametrize("block", [0xFF00, 3000])
def test_index_queries_match_whole_file(tmp_path, block):
    """Every query through the spec-built BAI returns exactly the records of
    the whole-file read that overlap the interval, in file order: intervals
    inside one 16 kb window, across windows and bins, before all reads,
    after all reads, start <= 0, one base, inside the 20 kb deletion only."""
    from pilon_amd.bam import IndexedBamFile
    refs, batches = spanning_batches()
    p = tmp_path / "q.bam"
    S.write_bam_indexed(str(p), refs, batches, block=block)
    with BamFile(p) as whole:
        wb = [whole.batch(0), whole.batch(1)]
    rng = np.random.default_rng(3)
    cases = [(0, 1, 100), (0, -5, 10), (0, 15_000, 16_000), (0, 1, 250_000), (0, 16_384, 16_384), (0, 245_000, 250_000),
             (0, 150_000, 150_000), (0, 90_000, 90_000), (1, 1, 60_000), (1, 55_000, 60_000), (0, 1000, 1000)]
    cases += [(int(r), int(a), int(a + rng.integers(0, 30_000))) for r, a in
              zip(rng.integers(0, 2, 40), rng.integers(1, 240_000, 40))]
    with IndexedBamFile(p) as ib:
        assert ib.names == ["chrA", "chrB"] and ib.lens == [250_000, 60_000]
        for ref, a, b in cases:
            q = ib.query(ref, a, b)
            same_reads(q, wb[ref], overlap_filter(wb[ref], a, b), (ref, a, b))
    # the deletion read is found inside its deletion (its bin spans 16 kb windows)
    with IndexedBamFile(p) as ib:
        assert 1000 in ib.query(0, 15_000, 15_001).pos.tolist()


def test_index_names_and_missing_index(tmp_path):
    """htsjdk's index names (<bam>.bai, <name>.bai) and Pilon's refusal of a
    BAM without an index (BamFile.scala:55-60)."""
    from pilon_amd.bam import IndexedBamFile
    refs, batches = spanning_batches()
    p = tmp_path / "x.bam"
    S.write_bam_indexed(str(p), refs, batches, index_path=str(tmp_path / "x.bai"))
    with IndexedBamFile(p) as ib:
        assert ib.query(1, 1, 60_000).n == batches[1].n
    (tmp_path / "x.bai").rename(tmp_path / "elsewhere.bai")
    with pytest.raises(BamError) as e:
        IndexedBamFile(p)
    assert e.value.status == 1 and "must be indexed" in str(e.value)
    with IndexedBamFile(p, index=tmp_path / "elsewhere.bai") as ib:
        assert ib.query(0, 1, 10).n >= 0
    bad = tmp_path / "bad.bai"
    bad.write_bytes(b"BAI\x01" + struct.pack("<i", 5))
    with pytest.raises(BamError):
        IndexedBamFile(p, index=bad)
    with BamFile(p) as whole, pytest.raises(BamError):
        lib_whole = whole
        from pilon_amd.bam import lib
        from pilon_amd.abi import PlReadSoa
        import ctypes as C
        s = PlReadSoa()
        st = lib().pl_bam_query(lib_whole.h, 0, 1, 10, C.byref(s))
        raise BamError(st, "query on a whole-file reader")


def test_index_query_synthetic_c2(tmp_path):
    """c2 x0.02 (clips, improper pairs, MAPQ-0 repeats, indels) through the
    index: random region queries equal the whole-file read filtered by
    overlap; a region's ±10 kb query gives the oracle the same region."""
    from oracle import oracle as O
    from pilon_amd.bam import IndexedBamFile
    from pilon_amd.synth import SynthData, preset
    d = SynthData(preset("c2", scale=0.02))
    b = d.batch(0)
    L = len(d.draft(0))
    p = tmp_path / "c2.bam"
    S.write_bam_indexed(str(p), [(d.names[0], L)], {0: b})
    with BamFile(p) as whole:
        wb = whole.batch(0)
    rng = np.random.default_rng(8)
    with IndexedBamFile(p) as ib:
        for a in rng.integers(1, L, 12):
            a = int(a)
            e = min(L, a + int(rng.integers(100, 30_000)))
            q = ib.query(0, a, e)
            same_reads(q, wb, overlap_filter(wb, a, e), (a, e))
        s, e = L // 3, L // 2
        q = ib.query(0, max(s - 10000, 0), min(e + 10000, L))
    cfg = default_config()
    r1 = O.run_region(cfg, d.draft(0), s, e, [q], name=d.names[0])
    r2 = O.run_region(cfg, d.draft(0), s, e, [b], name=d.names[0])
    assert r1["status"] == r2["status"] == 0
    assert np.array_equal(r1["pile"], r2["pile"]) and r1["changes"] == r2["changes"]
    assert r1["stats"] == r2["stats"]
