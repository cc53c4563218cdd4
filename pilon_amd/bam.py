"""BAM files -> read batches (libpilonbam.so, include/pilon_bam.h).

The reading half of Pilon's BamFile (BamFile.scala:38-60,108-148): the
sequence dictionary (getSeqs) and, per reference, every record placed on it
as a ``pl_read_soa``.  validateRead and the +-10 kb query of BamFile.process
are applied per region by the engine, so one batch per reference serves all
of its regions.
"""
import ctypes as C
import os

import numpy as np

from .abi import PlReadSoa, STATUS_NAMES
from .reads import ReadBatch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpilonbam.so")
_lib = None


class PlBamStats(C.Structure):
    _fields_ = [("compressed_bytes", C.c_int64), ("bam_bytes", C.c_int64), ("blocks", C.c_int64),
                ("records", C.c_int64), ("placed", C.c_int64), ("exotic", C.c_int64),
                ("inflate_s", C.c_double), ("decode_s", C.c_double), ("threads", C.c_int32),
                ("eof_marker", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BamError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB)
        P = C.POINTER
        sig = {
            "pl_bam_open": (C.c_int, [C.c_char_p, C.c_int32, P(C.c_void_p)]),
            "pl_bam_open_indexed": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int32, P(C.c_void_p)]),
            "pl_bam_query": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_int64, P(PlReadSoa)]),
            "pl_bam_n_refs": (C.c_int32, [C.c_void_p]),
            "pl_bam_ref_name": (C.c_char_p, [C.c_void_p, C.c_int32]),
            "pl_bam_ref_len": (C.c_int64, [C.c_void_p, C.c_int32]),
            "pl_bam_ref_reads": (C.c_int, [C.c_void_p, C.c_int32, P(PlReadSoa)]),
            "pl_bam_get_stats": (None, [C.c_void_p, P(PlBamStats)]),
            "pl_bam_stream_open": (C.c_int, [C.c_char_p, C.c_int32, P(C.c_void_p)]),
            "pl_bam_stream_bytes": (C.c_int64, [C.c_void_p]),
            "pl_bam_stream_next": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_int64), C.c_void_p, C.c_int64,
                                             P(C.c_int64)]),
            "pl_bam_close": (None, [C.c_void_p]),
            "pl_bam_last_error": (C.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class _RefReads:
    """One reference's records as a read set (library-owned arrays)."""

    def __init__(self, bam, ref):
        self.bam, self.ref = bam, ref

    def soa(self):
        return self.bam.soa(self.ref)


class BamFile:
    """A decoded coordinate-sorted BAM (BamFile.scala:38-60)."""

    def __init__(self, path, threads=0):
        h = C.c_void_p()
        st = lib().pl_bam_open(os.fsencode(path), threads, C.byref(h))
        if st:
            raise BamError(st, lib().pl_bam_last_error().decode())
        self.h = h
        n = lib().pl_bam_n_refs(h)
        self.names = [lib().pl_bam_ref_name(h, r).decode() for r in range(n)]
        self.lens = [lib().pl_bam_ref_len(h, r) for r in range(n)]

    def close(self):
        if getattr(self, "h", None):
            lib().pl_bam_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stats(self):
        s = PlBamStats()
        lib().pl_bam_get_stats(self.h, C.byref(s))
        return s.as_dict()

    def soa(self, ref):
        s = PlReadSoa()
        st = lib().pl_bam_ref_reads(self.h, ref, C.byref(s))
        if st:
            raise BamError(st, lib().pl_bam_last_error().decode())
        return s

    def reads(self, ref):
        """Reference `ref`'s reads for the engine or the oracle (valid while open)."""
        return _RefReads(self, ref)

    def batch(self, ref):
        """A numpy-owned copy of reference `ref`'s reads."""
        return _batch_of(self.soa(ref))


class IndexedBamFile(BamFile):
    """A coordinate-sorted BAM read through its BAI index: header and index
    at open, records per region (htsjdk's queryOverlapping, BamFile.scala:
    55-60,118-119)."""

    def __init__(self, path, index=None, threads=0):
        h = C.c_void_p()
        st = lib().pl_bam_open_indexed(os.fsencode(path), os.fsencode(index) if index else None, threads, C.byref(h))
        if st:
            raise BamError(st, lib().pl_bam_last_error().decode())
        self.h = h
        n = lib().pl_bam_n_refs(h)
        self.names = [lib().pl_bam_ref_name(h, r).decode() for r in range(n)]
        self.lens = [lib().pl_bam_ref_len(h, r) for r in range(n)]

    def query(self, ref, start, stop):
        """A numpy-owned copy of the reads of `ref` overlapping [start, stop] (1-based)."""
        s = PlReadSoa()
        st = lib().pl_bam_query(self.h, ref, int(start), int(stop), C.byref(s))
        if st:
            raise BamError(st, lib().pl_bam_last_error().decode())
        return _batch_of(s)


def _batch_of(s):
    """A numpy-owned copy of a library-owned read batch."""
    if True:
        n = s.n_reads

        def arr(ptr, count, ctype, dtype):
            if count == 0 or not ptr:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).astype(dtype, copy=True)

        co = arr(s.cigar_off, n + 1, C.c_int64, np.int64)
        so = arr(s.seq_off, n + 1, C.c_int64, np.int64)
        nb = int(so[-1])
        ex_r = arr(s.exo_read, s.n_exotic, C.c_int64, np.int64)
        ex_o = arr(s.exo_off, s.n_exotic, C.c_int64, np.int64)
        nib_len = 1
        if s.n_exotic:
            last = int(ex_r[-1])
            nib_len = (int(ex_o[-1]) + int(so[last + 1] - so[last]) + 1) // 2
        return ReadBatch(arr(s.pos, n, C.c_int32, np.int32), arr(s.tlen, n, C.c_int32, np.int32),
                         arr(s.flag, n, C.c_uint16, np.uint16), arr(s.mapq, n, C.c_uint8, np.uint8),
                         co, arr(s.cigar, int(co[-1]), C.c_uint32, np.uint32), so,
                         arr(s.bases2, (nb + 3) // 4, C.c_uint8, np.uint8),
                         arr(s.quals, nb, C.c_int8, np.int8), ex_r, ex_o,
                         arr(s.nib, nib_len, C.c_uint8, np.uint8))


class BamStream(BamFile):
    """A BAM read as slices of whole records (pl_bam_stream_open / _next):
    the header at open, then the record bytes, for a consumer that decodes
    them itself."""

    def __init__(self, path, threads=0):
        h = C.c_void_p()
        st = lib().pl_bam_stream_open(os.fsencode(path), threads, C.byref(h))
        if st:
            raise BamError(st, lib().pl_bam_last_error().decode())
        self.h = h
        n = lib().pl_bam_n_refs(h)
        self.names = [lib().pl_bam_ref_name(h, r).decode() for r in range(n)]
        self.lens = [lib().pl_bam_ref_len(h, r) for r in range(n)]

    @property
    def total_bytes(self):
        """The stream's inflated size, which bounds the record bytes it hands over."""
        return lib().pl_bam_stream_bytes(self.h)

    def next_into(self, dst, cap, rec_off, rec_cap):
        """Inflate the next slice into the buffer at address dst (cap bytes)
        and its record starts into rec_off (rec_cap int64s): (bytes, records),
        (0, 0) at the end."""
        ln, nr = C.c_int64(), C.c_int64()
        st = lib().pl_bam_stream_next(self.h, dst, cap, C.byref(ln), rec_off, rec_cap, C.byref(nr))
        if st:
            raise BamError(st, lib().pl_bam_last_error().decode())
        return ln.value, nr.value

    def slices(self, cap=64 << 20):
        """The slices as (numpy bytes, numpy record starts) pairs, into host
        buffers of this object's own."""
        import numpy as np
        buf = np.empty(cap, np.uint8)
        offs = np.empty(cap // 36 + 1, np.int64)
        while True:
            n, r = self.next_into(buf.ctypes.data, cap, offs.ctypes.data, len(offs))
            if n == 0:
                return
            yield buf[:n], offs[:r]
