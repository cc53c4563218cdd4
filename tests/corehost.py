"""Host build of the kernel core (libpl_core_host.so) — CPU-test helper only."""
import ctypes as C
import os

import numpy as np

from pilon_amd.abi import PlConfig, PlReadSoa, PlRegionStats, NFIELDS, call_dtype

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pilon_amd",
                   "libpl_core_host.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(LIB)
        L.ch_run.restype = C.c_void_p
        L.ch_run.argtypes = [C.POINTER(PlConfig), C.c_char_p, C.c_int64, C.c_char_p, C.POINTER(PlReadSoa),
                             C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.ch_text.restype = C.c_int64
        L.ch_text.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_char_p)]
        L.ch_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def run(cfg, contig, batch, regions, W=1024, name="ctg", lrt=None, set_kinds=1):
    """regions: list of (start, stop). Returns list of per-region dicts.
    lrt: longReadType per read (None: none); set_kinds: bit 0 a frags BAM
    among the set's, bit 1 a nanopore BAM."""
    L = lib()
    starts = np.array([r[0] for r in regions], np.int32)
    stops = np.array([r[1] for r in regions], np.int32)
    G = int((stops - starts + 1).sum())
    pile = np.zeros((NFIELDS, G), np.int32)
    calls = np.zeros(G, call_dtype())
    stats = (PlRegionStats * len(regions))()
    errs = np.zeros(len(regions), np.int32)
    soa = batch.soa()
    lr = None if lrt is None else np.ascontiguousarray(lrt, np.uint8)
    h = L.ch_run(C.byref(cfg), bytes(contig), len(contig), name.encode(), C.byref(soa), len(regions),
                 starts.ctypes.data, stops.ctypes.data, W, pile.ctypes.data, calls.ctypes.data,
                 C.addressof(stats), errs.ctypes.data,
                 None if lr is None else lr.ctypes.data, set_kinds)
    out = []
    off = 0
    for j, (s, e) in enumerate(regions):
        n = e - s + 1
        p = C.c_char_p()
        ln = L.ch_text(h, j, 0, C.byref(p))
        changes = C.string_at(p, ln).decode()
        ln = L.ch_text(h, j, 1, C.byref(p))
        bases = C.string_at(p, ln)
        out.append(dict(status=int(errs[j]), pile=pile[:, off:off + n], calls=calls[off:off + n],
                        stats=stats[j].as_dict(), changes=changes, bases=bases))
        off += n
    L.ch_free(h)
    return out
