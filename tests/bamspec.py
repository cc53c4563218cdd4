"""A BAM writer written literally from the SAMv1 specification (sections
4.1 BGZF and 4.2 BAM), used as the independent encoder in the BAM reader's
tests: records are packed field by field with struct, SEQ two 4-bit codes
per byte with the first base in the high nibble, QUAL 0xFF when absent."""
import struct
import zlib

import numpy as np

from pilon_amd.abi import PL_RF_EXOTIC, PL_RF_MATE_SAME_REF, PL_RF_NO_QUAL
from pilon_amd.reads import NIB_CODES, unpack_2bit

# SAMv1 section 4.1.2, the end-of-file marker, byte for byte
SPEC_EOF = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
NIB = {c: i for i, c in enumerate(NIB_CODES)}


def bgzf_block(data, level=6):
    """One BGZF block: gzip header with the BC extra subfield, raw deflate, CRC32, ISIZE."""
    c = zlib.compressobj(level, zlib.DEFLATED, -15)
    cdata = c.compress(data) + c.flush()
    bsize = 18 + len(cdata) + 8
    head = struct.pack("<BBBBIBBHBBHH", 31, 139, 8, 4, 0, 0, 255, 6, 66, 67, 2, bsize - 1)
    return head + cdata + struct.pack("<II", zlib.crc32(data) & 0xFFFFFFFF, len(data))


def bgzf(data, block=0xFF00, level=6, eof=True):
    out = bytearray()
    for i in range(0, len(data), block):
        out += bgzf_block(data[i:i + block], level)
    if eof:
        out += bgzf_block(b"")
    return bytes(out)


def reg2bin(beg, end):
    """SAMv1 section 5.3: the smallest bin holding [beg, end)."""
    end -= 1
    if beg >> 14 == end >> 14:
        return ((1 << 15) - 1) // 7 + (beg >> 14)
    if beg >> 17 == end >> 17:
        return ((1 << 12) - 1) // 7 + (beg >> 17)
    if beg >> 20 == end >> 20:
        return ((1 << 9) - 1) // 7 + (beg >> 20)
    if beg >> 23 == end >> 23:
        return ((1 << 6) - 1) // 7 + (beg >> 23)
    if beg >> 26 == end >> 26:
        return ((1 << 3) - 1) // 7 + (beg >> 26)
    return 0


def header(refs, text=b"@HD\tVN:1.6\tSO:coordinate\n"):
    out = bytearray(b"BAM\x01") + struct.pack("<i", len(text)) + text + struct.pack("<i", len(refs))
    for name, ln in refs:
        nb = name.encode() + b"\0"
        out += struct.pack("<i", len(nb)) + nb + struct.pack("<i", ln)
    return bytes(out)


def record(ref, pos0, cigar, seq, qual, flag=0, mapq=60, next_ref=-1, next_pos=-1, tlen=0, name="r", tags=b""):
    """One alignment record.  seq: a string over "=ACMGRSVTWYHKDBN" ("" = *);
    qual: a list of ints or None (absent: 0xFF bytes); cigar: BAM words."""
    rlen = sum(w >> 4 for w in cigar if (w & 15) in (0, 2, 3, 7, 8))
    nm = name.encode() + b"\0"
    codes = [NIB[c] for c in seq]
    if len(codes) % 2:
        codes.append(0)
    packed = bytes((codes[k] << 4) | codes[k + 1] for k in range(0, len(codes), 2))
    q = bytes([0xFF] * len(seq)) if qual is None else bytes(v & 0xFF for v in qual)
    body = struct.pack("<iiBBHHHIiii", ref, pos0, len(nm), mapq, reg2bin(max(pos0, 0), max(pos0, 0) + max(rlen, 1)),
                       len(cigar), flag, len(seq), next_ref, next_pos, tlen)
    body += nm + struct.pack(f"<{len(cigar)}I", *cigar) + packed + q + tags
    return struct.pack("<I", len(body)) + body


def batch_records(batch, ref):
    """The records a ReadBatch of reference `ref` stands for (flags without
    the PL_RF_* bits; a mate on the same reference iff PL_RF_MATE_SAME_REF)."""
    out = []
    so, co = batch.seq_off, batch.cigar_off
    exo = {int(r): int(o) for r, o in zip(batch.exo_read, batch.exo_off)}
    for i in range(batch.n):
        s0, s1 = int(so[i]), int(so[i + 1])
        L = s1 - s0
        f = int(batch.flag[i])
        if i in exo:
            o = exo[i]
            seq = "".join(NIB_CODES[(batch.nib[(o + k) >> 1] >> (4 * (1 - ((o + k) & 1)))) & 15] for k in range(L))
        else:
            seq = "".join("ACGT"[c] for c in unpack_2bit(batch.bases2, s0, L))
        qual = None if f & PL_RF_NO_QUAL else [int(v) for v in batch.quals[s0:s1]]
        out.append(record(ref, int(batch.pos[i]) - 1, [int(w) for w in batch.cigar[int(co[i]):int(co[i + 1])]], seq,
                          qual, flag=f & 0x0FFF, mapq=int(batch.mapq[i]),
                          next_ref=ref if f & PL_RF_MATE_SAME_REF else -1, next_pos=-1,
                          tlen=int(batch.tlen[i]), name=f"r{i}"))
    return out


def write_bam(path, refs, batches, block=0xFF00, level=1, eof=True):
    """refs: [(name, length)]; batches: {refID: ReadBatch} written in refID order."""
    data = bytearray(header(refs))
    for ref in sorted(batches):
        for rec in batch_records(batches[ref], ref):
            data += rec
    with open(path, "wb") as f:
        f.write(bgzf(bytes(data), block=block, level=level, eof=eof))
    return len(data)


def same_batch(a, b, what=""):
    """Field-by-field equality of two read batches: quals of reads that have
    them, 2-bit bases of ACGT reads, nibbles of the others."""
    n = a.n
    assert b.n == n, (what, a.n, b.n)
    for k in ("pos", "tlen", "flag", "mapq", "cigar_off", "seq_off"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), (what, k)
    assert np.array_equal(a.cigar[:int(a.cigar_off[-1])], b.cigar[:int(b.cigar_off[-1])]), (what, "cigar")
    nb = int(a.seq_off[-1])
    lens = np.diff(a.seq_off)
    has_q = np.repeat((a.flag & PL_RF_NO_QUAL) == 0, lens)
    assert np.array_equal(a.quals[:nb][has_q], b.quals[:nb][has_q]), (what, "quals")
    acgt = np.repeat((a.flag & PL_RF_EXOTIC) == 0, lens)
    assert np.array_equal(unpack_2bit(a.bases2, 0, nb)[acgt], unpack_2bit(b.bases2, 0, nb)[acgt]), (what, "bases")
    assert np.array_equal(a.exo_read, b.exo_read), (what, "exo_read")
    for r, oa, ob in zip(a.exo_read, a.exo_off, b.exo_off):
        L = int(a.seq_off[r + 1] - a.seq_off[r])
        k = np.arange(L)
        na = (a.nib[(oa + k) >> 1] >> (4 * (1 - ((oa + k) & 1)))) & 15
        nb_ = (b.nib[(ob + k) >> 1] >> (4 * (1 - ((ob + k) & 1)))) & 15
        assert np.array_equal(na, nb_), (what, "nib", int(r))


def bgzf_offsets(data_len, block, level_blocks):
    """Compressed offset of each BGZF block of a stream cut every `block`
    bytes (the end-of-file block last), from the blocks' sizes."""
    coff, at = [], 0
    for b in level_blocks:
        coff.append(at)
        at += len(b)
    return coff


def write_bam_indexed(path, refs, batches, block=0xFF00, level=1, index_path=None, meta_bin=True):
    """write_bam and, beside it, the BAI index built from SAMv1 section 5.2:
    per reference the chunks [virtual start, virtual end) of consecutive
    records of each bin (virtual offset = compressed block offset << 16 |
    offset in the inflated block), the linear index of 16 kb windows (the
    smallest virtual offset of a record overlapping the window, 0 for a
    window without one), and the 37450 pseudo-bin of counts."""
    data = bytearray(header(refs))
    recs = []                                   # (ref, pos0, end0, u_start, u_end, mapped)
    for ref in sorted(batches):
        for rec in batch_records(batches[ref], ref):
            pos0 = struct.unpack_from("<i", rec, 8)[0]
            ncig = struct.unpack_from("<H", rec, 16)[0]
            flag = struct.unpack_from("<H", rec, 18)[0]
            l_name = rec[12]
            cig = struct.unpack_from(f"<{ncig}I", rec, 36 + l_name)
            rlen = sum(w >> 4 for w in cig if (w & 15) in (0, 2, 3, 7, 8))
            recs.append((ref, pos0, pos0 + max(rlen, 1), len(data), len(data) + len(rec), not flag & 4))
            data += rec
    blocks = [bgzf_block(bytes(data[i:i + block]), level) for i in range(0, len(data), block)]
    blocks.append(bgzf_block(b""))
    coff = bgzf_offsets(len(data), block, blocks)

    def voff(u):
        return (coff[u // block] << 16) | (u % block)

    with open(path, "wb") as f:
        for b in blocks:
            f.write(b)
    out = bytearray(b"BAI\x01") + struct.pack("<i", len(refs))
    for ref in range(len(refs)):
        mine = [r for r in recs if r[0] == ref]
        bins = {}
        lin = {}
        for _, p0, e0, u0, u1, _m in mine:
            b = reg2bin(p0, e0)
            ch = bins.setdefault(b, [])
            if ch and ch[-1][1] == voff(u0):
                ch[-1][1] = voff(u1)
            else:
                ch.append([voff(u0), voff(u1)])
            for w in range(p0 >> 14, ((e0 - 1) >> 14) + 1):
                lin.setdefault(w, voff(u0))
        items = sorted(bins.items())
        if meta_bin and mine:
            nm = sum(1 for r in mine if r[5])
            items.append((37450, [[voff(mine[0][3]), voff(mine[-1][4])], [nm, len(mine) - nm]]))
        out += struct.pack("<i", len(items))
        for b, ch in items:
            out += struct.pack("<Ii", b, len(ch))
            for a, e in ch:
                out += struct.pack("<QQ", a, e)
        n_intv = max(lin) + 1 if lin else 0
        out += struct.pack("<i", n_intv)
        for w in range(n_intv):
            out += struct.pack("<Q", lin.get(w, 0))
    out += struct.pack("<Q", 0)                  # n_no_coor
    with open(index_path or path + ".bai", "wb") as f:
        f.write(bytes(out))
    return len(data)
