"""Host-side packer: SAM-like reads -> the packed structure-of-arrays the C ABI takes.

This is the part of a JVM/C++ host that turns htsjdk SAMRecords into
``pl_read_soa`` (include/pilon_hip.h).  The fields it packs are exactly the
SAMRecord getters PileUpRegion.addRead reads (PileUpRegion.scala:103-136):
alignment start, inferred insert size, flags (plus the mate-on-same-contig
bit used by the validity test at :107), MAPQ, CIGAR, bases and qualities.
"""
import re
import numpy as np

from .abi import PlReadSoa, PL_RF_NO_QUAL, PL_RF_MATE_SAME_REF, PL_RF_EXOTIC

CIGAR_OPS = "MIDNSHP=X"
OP_CODE = {c: i for i, c in enumerate(CIGAR_OPS)}
NIB_CODES = "=ACMGRSVTWYHKDBN"
NIB_OF = {c: i for i, c in enumerate(NIB_CODES)}
TWO_BIT = {"A": 0, "C": 1, "G": 2, "T": 3}
_CIGAR_RE = re.compile(r"(\d+)([MIDNSHP=X])")


def parse_cigar(cigar):
    """'10S90M' -> BAM words (len << 4 | op)."""
    if cigar in ("*", ""):
        return []
    parts = _CIGAR_RE.findall(cigar)
    if "".join(n + o for n, o in parts) != cigar:
        raise ValueError(f"bad CIGAR {cigar!r}")
    return [(int(n) << 4) | OP_CODE[o] for n, o in parts]


def cigar_ref_len(words):
    return sum(w >> 4 for w in words if (w & 0xF) in (0, 2, 3, 7, 8))


def cigar_read_len(words):
    return sum(w >> 4 for w in words if (w & 0xF) in (0, 1, 4, 7, 8))


class ReadBatch:
    """A sorted batch of one contig's reads in pl_read_soa layout (numpy-owned)."""

    def __init__(self, pos, tlen, flag, mapq, cigar_off, cigar, seq_off, bases2, quals,
                 exo_read=None, exo_off=None, nib=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.int32)
        self.tlen = np.ascontiguousarray(tlen, dtype=np.int32)
        self.flag = np.ascontiguousarray(flag, dtype=np.uint16)
        self.mapq = np.ascontiguousarray(mapq, dtype=np.uint8)
        self.cigar_off = np.ascontiguousarray(cigar_off, dtype=np.int64)
        self.cigar = np.ascontiguousarray(cigar, dtype=np.uint32)
        self.seq_off = np.ascontiguousarray(seq_off, dtype=np.int64)
        self.bases2 = np.ascontiguousarray(bases2, dtype=np.uint8)
        self.quals = np.ascontiguousarray(quals, dtype=np.int8)
        self.exo_read = np.ascontiguousarray(exo_read if exo_read is not None else [], dtype=np.int64)
        self.exo_off = np.ascontiguousarray(exo_off if exo_off is not None else [], dtype=np.int64)
        self.nib = np.ascontiguousarray(nib if nib is not None else np.zeros(1, np.uint8), dtype=np.uint8)
        n = len(self.pos)
        if not (len(self.tlen) == len(self.flag) == len(self.mapq) == n):
            raise ValueError("header arrays differ in length")
        if len(self.cigar_off) != n + 1 or len(self.seq_off) != n + 1:
            raise ValueError("offset arrays must have n+1 entries")
        if n > 1 and np.any(np.diff(self.pos) < 0):
            raise ValueError("reads must be coordinate-sorted")
        # keep the 2-bit stream readable a few bytes past its end
        if len(self.bases2) < (int(self.seq_off[-1]) + 3) // 4 + 8:
            b = np.zeros((int(self.seq_off[-1]) + 3) // 4 + 8, np.uint8)
            b[:len(self.bases2)] = self.bases2
            self.bases2 = b
        if len(self.quals) < int(self.seq_off[-1]) + 8:
            q = np.zeros(int(self.seq_off[-1]) + 8, np.int8)
            q[:len(self.quals)] = self.quals
            self.quals = q

    @property
    def n(self):
        return len(self.pos)

    def soa(self):
        """ctypes pl_read_soa pointing into this batch (keep the batch alive)."""
        s = PlReadSoa()
        s.n_reads = self.n
        s.pos = self.pos.ctypes.data
        s.tlen = self.tlen.ctypes.data
        s.flag = self.flag.ctypes.data
        s.mapq = self.mapq.ctypes.data
        s.cigar_off = self.cigar_off.ctypes.data
        s.cigar = self.cigar.ctypes.data if len(self.cigar) else self.cigar_off.ctypes.data
        s.seq_off = self.seq_off.ctypes.data
        s.bases2 = self.bases2.ctypes.data
        s.quals = self.quals.ctypes.data
        s.n_exotic = len(self.exo_read)
        s.exo_read = self.exo_read.ctypes.data if len(self.exo_read) else 0
        s.exo_off = self.exo_off.ctypes.data if len(self.exo_off) else 0
        s.nib = self.nib.ctypes.data
        return s

    def slice(self, lo, hi):
        """Reads [lo, hi) as a new batch (offsets rebased)."""
        co, so = self.cigar_off, self.seq_off
        c0, c1, s0, s1 = int(co[lo]), int(co[hi]), int(so[lo]), int(so[hi])
        # 2-bit stream: copy base by base window via unpack for simplicity
        codes = unpack_2bit(self.bases2, s0, s1 - s0)
        ex = (self.exo_read >= lo) & (self.exo_read < hi)
        nib_codes = []
        exo_read, exo_off = [], []
        for k in np.nonzero(ex)[0]:
            r = int(self.exo_read[k])
            L = int(so[r + 1] - so[r])
            exo_read.append(r - lo)
            exo_off.append(len(nib_codes))
            nib_codes.extend(unpack_nib(self.nib, int(self.exo_off[k]), L))
        return ReadBatch(self.pos[lo:hi], self.tlen[lo:hi], self.flag[lo:hi], self.mapq[lo:hi],
                         co[lo:hi + 1] - c0, self.cigar[c0:c1], so[lo:hi + 1] - s0,
                         pack_2bit(codes), self.quals[s0:s1], exo_read, exo_off,
                         pack_nib(nib_codes))

    def take(self, idx):
        """The reads at the increasing indices idx as a new batch (a subset
        keeps the coordinate order)."""
        idx = np.asarray(idx, dtype=np.int64)
        if len(idx) > 1 and np.any(np.diff(idx) <= 0):
            raise ValueError("indices must be increasing")
        co, so = self.cigar_off, self.seq_off

        def gather(off, idx):
            ln = off[idx + 1] - off[idx]
            noff = np.concatenate([[0], np.cumsum(ln)]).astype(np.int64)
            src = np.repeat(off[idx] - noff[:-1], ln) + np.arange(int(noff[-1]), dtype=np.int64)
            return noff, src

        nco, csrc = gather(co, idx)
        nso, ssrc = gather(so, idx)
        codes = unpack_2bit(self.bases2, 0, int(so[-1]))[ssrc] if len(ssrc) else np.zeros(0, np.uint8)
        where = {int(r): k for k, r in enumerate(self.exo_read)}
        exo_read, exo_off, nib_codes = [], [], []
        for j, r in enumerate(idx):
            k = where.get(int(r))
            if k is not None:
                L = int(so[r + 1] - so[r])
                exo_read.append(j)
                exo_off.append(len(nib_codes))
                nib_codes.extend(unpack_nib(self.nib, int(self.exo_off[k]), L))
        return ReadBatch(self.pos[idx], self.tlen[idx], self.flag[idx], self.mapq[idx], nco, self.cigar[csrc], nso,
                         pack_2bit(codes), self.quals[ssrc], exo_read, exo_off, pack_nib(nib_codes))


def pack_2bit(codes):
    codes = np.asarray(codes, dtype=np.uint8)
    n = len(codes)
    pad = (-n) % 4
    c = np.concatenate([codes, np.zeros(pad, np.uint8)]).reshape(-1, 4)
    out = (c[:, 0] | (c[:, 1] << 2) | (c[:, 2] << 4) | (c[:, 3] << 6)).astype(np.uint8)
    return out


def unpack_2bit(packed, start, n):
    idx = np.arange(start, start + n, dtype=np.int64)
    return ((packed[idx >> 2] >> (2 * (idx & 3)).astype(np.uint8)) & 3).astype(np.uint8)


def pack_nib(codes):
    codes = np.asarray(codes, dtype=np.uint8)
    if len(codes) % 2:
        codes = np.concatenate([codes, np.zeros(1, np.uint8)])
    if len(codes) == 0:
        return np.zeros(1, np.uint8)
    c = codes.reshape(-1, 2)
    return ((c[:, 0] << 4) | c[:, 1]).astype(np.uint8)


def unpack_nib(packed, start, n):
    idx = np.arange(start, start + n, dtype=np.int64)
    byte = packed[idx >> 1]
    return np.where(idx & 1, byte & 0xF, byte >> 4).astype(np.uint8)


def pack_reads(reads, sort=True):
    """Pack an iterable of SAM-like reads into a ReadBatch.

    Each read is a dict with keys pos (1-based), cigar (str), seq (str, '*' or
    '' for none), qual (list of ints, or None for absent), flag, mapq, tlen and
    optionally mate_same_ref (default True).
    """
    reads = list(reads)
    if sort:
        reads = sorted(reads, key=lambda r: r["pos"])
    n = len(reads)
    pos = np.zeros(n, np.int32)
    tlen = np.zeros(n, np.int32)
    flag = np.zeros(n, np.uint16)
    mapq = np.zeros(n, np.uint8)
    cigar_off = np.zeros(n + 1, np.int64)
    seq_off = np.zeros(n + 1, np.int64)
    cigar, codes, quals = [], [], []
    exo_read, exo_off, nib_codes = [], [], []
    for i, r in enumerate(reads):
        words = parse_cigar(r.get("cigar", "*"))
        seq = r.get("seq", "*")
        if seq == "*":
            seq = ""
        q = r.get("qual")
        f = int(r.get("flag", 0)) & 0x0FFF
        if r.get("mate_same_ref", True):
            f |= PL_RF_MATE_SAME_REF
        if q is None:
            f |= PL_RF_NO_QUAL
            q = [0] * len(seq)
        if len(q) != len(seq):
            raise ValueError("qual and seq lengths differ")
        if any(c not in TWO_BIT for c in seq):
            f |= PL_RF_EXOTIC
            exo_read.append(i)
            exo_off.append(len(nib_codes))
            nib_codes.extend(NIB_OF[c] for c in seq)
            codes.extend([0] * len(seq))
        else:
            codes.extend(TWO_BIT[c] for c in seq)
        pos[i] = r["pos"]
        tlen[i] = r.get("tlen", 0)
        flag[i] = f
        mapq[i] = r.get("mapq", 60)
        cigar.extend(words)
        quals.extend(q)
        cigar_off[i + 1] = len(cigar)
        seq_off[i + 1] = len(codes)
    return ReadBatch(pos, tlen, flag, mapq, cigar_off, np.array(cigar, np.uint32), seq_off,
                     pack_2bit(codes), np.array(quals, np.int64).astype(np.int8), exo_read, exo_off,
                     pack_nib(nib_codes))


def aligned_bases_in_region(batch, start, stop):
    """M/=/X bases whose locus is in [start, stop] — the throughput unit."""
    total = 0
    for i in range(batch.n):
        loc = int(batch.pos[i])
        for w in batch.cigar[batch.cigar_off[i]:batch.cigar_off[i + 1]]:
            op, ln = int(w) & 0xF, int(w) >> 4
            if op in (0, 7, 8):
                a, b = max(loc, start), min(loc + ln - 1, stop)
                if b >= a:
                    total += b - a + 1
            if op in (0, 2, 3, 7, 8):
                loc += ln
    return total
