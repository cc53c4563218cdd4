"""Python handle on the C oracle (pilon_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product (pilon_amd/).
"""
import ctypes as C
import re
import os
import subprocess

import numpy as np

from pilon_amd.abi import PlConfig, PlReadSoa, PlCall, PlRegionStats, NFIELDS, call_dtype

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpilon_oracle.so")
SRC = os.path.join(HERE, "pilon_oracle.c")


def build():
    """Compile the oracle with gcc (oracle/Makefile does the same)."""
    subprocess.check_call(["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-Wall", "-Wno-unused-function",
                           "-o", LIB, SRC, "-lm"])


class OrcResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("edge_read", C.c_int64), ("edge_why", C.c_char * 64),
                ("size", C.c_int32), ("pile", C.POINTER(C.c_int32)), ("bcq", C.POINTER(C.c_int64)),
                ("calls", C.POINTER(PlCall)), ("coverage", C.POINTER(C.c_int32)),
                ("badCoverage", C.POINTER(C.c_int32)), ("physCoverage", C.POINTER(C.c_int32)),
                ("insertSize", C.POINTER(C.c_int32)), ("weightedQual", C.POINTER(C.c_int8)),
                ("weightedMq", C.POINTER(C.c_int8)), ("clips", C.POINTER(C.c_int16)),
                ("stats", PlRegionStats), ("changes_text", C.c_char_p), ("changes_len", C.c_int64),
                ("vcf_text", C.c_void_p), ("vcf_len", C.c_int64), ("bed_text", C.c_void_p), ("bed_len", C.c_int64),
                ("tracks", C.POINTER(C.c_int32)),
                ("bases", C.POINTER(C.c_uint8)), ("bases_len", C.c_int64),
                ("n_indel", C.c_int32), ("indel_pos", C.POINTER(C.c_int32)),
                ("indel_off", C.POINTER(C.c_int64)), ("indel_str", C.POINTER(C.c_uint8))]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        L = C.CDLL(LIB)
        L.orc_region_run.restype = C.POINTER(OrcResult)
        L.orc_region_run.argtypes = [C.POINTER(PlConfig), C.c_char_p, C.c_int64, C.c_char_p,
                                     C.c_int32, C.c_int32, C.POINTER(C.POINTER(PlReadSoa)),
                                     C.c_int32, C.c_char_p, C.c_int64, C.c_int32]
        L.orc_region_run_typed.restype = C.POINTER(OrcResult)
        L.orc_region_run_typed.argtypes = [C.POINTER(PlConfig), C.c_char_p, C.c_int64, C.c_char_p,
                                           C.c_int32, C.c_int32, C.POINTER(C.POINTER(PlReadSoa)),
                                           C.c_void_p, C.c_int32, C.c_char_p, C.c_int64, C.c_int32]
        L.orc_region_run_big.restype = C.POINTER(OrcResult)
        L.orc_region_run_big.argtypes = [C.POINTER(PlConfig), C.c_char_p, C.c_int64, C.c_char_p,
                                         C.c_int32, C.c_int32, C.POINTER(C.POINTER(PlReadSoa)),
                                         C.c_void_p, C.c_int32, C.c_char_p, C.c_int64, C.c_int32,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.orc_result_free.argtypes = [C.POINTER(OrcResult)]
        L.orc_fmt_af.argtypes = [C.c_float, C.c_char_p]
        L.orc_smooth.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]
        L.orc_gc.argtypes = [C.c_char_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p]
        L.orc_kmer_graph.restype = C.c_int32
        L.orc_kmer_graph.argtypes = [C.POINTER(PlConfig), C.POINTER(PlReadSoa), C.c_int32, C.c_int32,
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_int64)]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_summary_regions.restype = C.c_int32
        L.orc_summary_regions.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_int32]
        L.orc_java_round.restype = C.c_int64
        L.orc_java_round.argtypes = [C.c_double]
        L.orc_roundDivI.restype = C.c_int32
        L.orc_roundDivI.argtypes = [C.c_int32, C.c_int32]
        L.orc_roundDivL.restype = C.c_int64
        L.orc_roundDivL.argtypes = [C.c_int64, C.c_int64]
        L.orc_basecall.restype = None
        L.orc_basecall.argtypes = [C.POINTER(PlConfig), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                   C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_int32, C.POINTER(C.c_int32), C.c_char_p,
                                   C.c_int32, C.POINTER(C.c_int32), C.c_char_p,
                                   C.POINTER(PlCall), C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


VCF_ON, VCF_DIPLOID, VCF_QE, TRACKS, BED = 1, 2, 4, 8, 16
TRACK_NAMES = ["Changes", "Unconfirmed", "Copy Number", "Coverage", "Bad Coverage", "Pct Bad", "Delta Coverage",
               "Dip Coverage", "Physical Coverage", "Clipped Alignments", "Weighted Qual", "Weighted MQ", "GC"]


def run_region(cfg, contig, start, stop, readsets, name="ctg", new_name=None, offset=0, vcf=0, set_types=None,
               big=None):
    """One GenomeRegion through the oracle; returns a dict of numpy arrays.
    vcf: VCF_ON (| VCF_DIPLOID | VCF_QE) adds the region's VCF record lines
    (GenomeRegion.writeVcf) as out["vcf"]; TRACKS the 13 per-locus tracks as
    out["tracks"]; BED the region's Pilon.bed lines as out["bed"].  set_types:
    one per read set, 1 for a "jumps" BAM (left out of fragment coverage).
    big: bigFixList as (locus, was, patch) tuples in List order (local
    reassembly's fixes, the last found first)."""
    L = lib()
    if isinstance(contig, str):
        contig = contig.encode()
    contig = bytes(contig)
    if new_name is None:
        new_name = name + "_pilon"
    soas = [rs.soa() for rs in readsets]
    arr = (C.POINTER(PlReadSoa) * max(1, len(soas)))(*[C.pointer(s) for s in soas])
    types = None
    if set_types is not None:
        types = (C.c_int32 * max(1, len(set_types)))(*set_types)
    big = list(big or [])
    nb = len(big)
    b_loc = (C.c_int32 * max(1, nb))(*[f[0] for f in big])
    b_was = (C.c_char_p * max(1, nb))(*[f[1].encode() for f in big])
    b_patch = (C.c_char_p * max(1, nb))(*[f[2].encode() for f in big])
    rp = L.orc_region_run_big(C.byref(cfg), contig, len(contig), name.encode(), start, stop,
                              arr, C.cast(types, C.c_void_p) if types is not None else None, len(soas),
                              new_name.encode(), offset, vcf, C.cast(b_loc, C.c_void_p), C.cast(b_was, C.c_void_p),
                              C.cast(b_patch, C.c_void_p), nb)
    r = rp.contents
    out = {"status": r.status, "edge_read": r.edge_read, "edge_why": r.edge_why.decode()}
    if r.status == 0:
        size = r.size
        out["size"] = size
        out["pile"] = _arr(r.pile, NFIELDS * size, np.int32).reshape(NFIELDS, size)
        out["bcq"] = _arr(r.bcq, 8 * size, np.int64).reshape(8, size)
        raw = np.ctypeslib.as_array(C.cast(r.calls, C.POINTER(C.c_uint8)), shape=(8 * size,)).copy()
        out["calls"] = raw.view(call_dtype())
        out["coverage"] = _arr(r.coverage, size, np.int32)
        out["badCoverage"] = _arr(r.badCoverage, size, np.int32)
        out["physCoverage"] = _arr(r.physCoverage, size, np.int32)
        out["insertSize"] = _arr(r.insertSize, size, np.int32)
        out["weightedQual"] = _arr(r.weightedQual, size, np.int8)
        out["weightedMq"] = _arr(r.weightedMq, size, np.int8)
        out["clips"] = _arr(r.clips, size, np.int16)
        out["stats"] = r.stats.as_dict()
        out["changes"] = C.string_at(r.changes_text, r.changes_len).decode() if r.changes_len else ""
        if vcf & VCF_ON:
            out["vcf"] = C.string_at(r.vcf_text, r.vcf_len).decode() if r.vcf_len else ""
        if vcf & BED:
            out["bed"] = C.string_at(r.bed_text, r.bed_len).decode() if r.bed_len else ""
        if vcf & (TRACKS | BED):
            out["tracks"] = _arr(r.tracks, len(TRACK_NAMES) * size, np.int32).reshape(len(TRACK_NAMES), size)
        out["bases"] = bytes(_arr(r.bases, r.bases_len, np.uint8))
        n = r.n_indel
        pos = _arr(r.indel_pos, n, np.int32)
        off = _arr(r.indel_off, n + 1, np.int64)
        sb = bytes(_arr(r.indel_str, int(off[-1]) if n else 0, np.uint8))
        out["indels"] = {int(pos[k]): sb[off[k]:off[k + 1]].decode() for k in range(n)}
    L.orc_result_free(rp)
    return out


def run_region_stats(cfg, contig, start, stop, readsets, name="ctg"):
    """run_region without copying the per-position results out: (status, stats)."""
    L = lib()
    contig = contig.encode() if isinstance(contig, str) else bytes(contig)
    soas = [rs.soa() for rs in readsets]
    arr = (C.POINTER(PlReadSoa) * max(1, len(soas)))(*[C.pointer(s) for s in soas])
    rp = L.orc_region_run(C.byref(cfg), contig, len(contig), name.encode(), start, stop,
                          arr, len(soas), (name + "_pilon").encode(), 0, 0)
    r = rp.contents
    status, stats = r.status, (r.stats.as_dict() if r.status == 0 else None)
    L.orc_result_free(rp)
    return status, stats


def fmt_af(af):
    """Java's "%.2f".format of a Float widened to Double, as the oracle restates it."""
    buf = C.create_string_buffer(32)
    lib().orc_fmt_af(C.c_float(af), buf)
    return buf.value.decode()


def smooth(values, window=200):
    v = np.ascontiguousarray(values, dtype=np.int32)
    out = np.zeros(len(v), np.int32)
    lib().orc_smooth(v.ctypes.data, len(v), window, out.ctypes.data)
    return out


def gc_track(contig, start, stop):
    contig = contig.encode() if isinstance(contig, str) else bytes(contig)
    out = np.zeros(stop - start + 1, np.int8)
    lib().orc_gc(contig, len(contig), start, stop, out.ctypes.data)
    return out


def java_round(x):
    return lib().orc_java_round(float(x))


def basecall(cfg, bc, qs, mqSum=0, qSum=0, deletions=0, insertions=0, insQual=0, delQual=0,
             ins_list=(), del_list=()):
    """BaseCall on a hand-made PileUp; returns (pl_call dict, indel_len, score)."""
    L = lib()
    b = (C.c_int64 * 4)(*bc)
    q = (C.c_int64 * 4)(*qs)
    il = (C.c_int32 * max(1, len(ins_list)))(*[len(s) for s in ins_list])
    dl = (C.c_int32 * max(1, len(del_list)))(*[len(s) for s in del_list])
    out = PlCall()
    ilen = C.c_int32()
    score = C.c_int64()
    L.orc_basecall(C.byref(cfg), b, q, mqSum, qSum, deletions, insertions, insQual, delQual,
                   len(ins_list), il, "".join(ins_list).encode(), len(del_list), dl,
                   "".join(del_list).encode(), C.byref(out), C.byref(ilen), C.byref(score))
    return {"base": out.base, "alt": out.alt, "flags": out.flags, "q": out.q}, ilen.value, score.value


def summary_regions(test, start, slop=100):
    """GenomeRegion.summaryRegions over a boolean array of a region starting
    at locus `start`: [(start, stop)] loci, nearEdge regions dropped."""
    t = np.ascontiguousarray(np.asarray(test, dtype=np.uint8))
    cap = len(t) + 1
    out = np.zeros(2 * cap, np.int32)
    n = lib().orc_summary_regions(t.ctypes.data, len(t), start, slop, out.ctypes.data, cap)
    return [(int(out[2 * k]), int(out[2 * k + 1])) for k in range(n)]


def kmer_graph(cfg, batch, K=47, min_depth=5):
    """Assembler.addReads + buildGraph over one read batch: (status, [(kmer,
    kGraph base, altGraph base or None)] in byte order of the K-mers)."""
    L = lib()
    soa = batch.soa()
    k, nx, al = C.c_void_p(), C.c_void_p(), C.c_void_p()
    n = C.c_int64()
    st = L.orc_kmer_graph(C.byref(cfg), C.byref(soa), K, min_depth, C.byref(k), C.byref(nx), C.byref(al),
                          C.byref(n))
    out = []
    if st == 0 and n.value:
        ks = C.string_at(k, K * n.value)
        ns = C.string_at(nx, n.value)
        als = C.string_at(al, n.value)
        out = [(ks[K * i:K * i + K].decode(), chr(ns[i]), chr(als[i]) if als[i] else None) for i in range(n.value)]
    for p in (k, nx, al):
        if p.value:
            L.orc_free(p)
    return st, out


# ---------------------------------------------------------------- GenomeFile
def _jint(x):
    """Java Int wrap."""
    return (x + 2 ** 31) % 2 ** 32 - 2 ** 31


def contig_regions(length, chunk_size=10000000):
    """GenomeFile.contigRegions (GenomeFile.scala:67-74), Int arithmetic:
    [(start, stop)] of Range(1, len + 1, chunkSize)."""
    n_chunks = int(_jint(length + chunk_size - 1) / chunk_size)      # Java division truncates
    chunk = int(_jint(length + n_chunks - 1) / n_chunks)
    return [(b, min(length, b + chunk - 1)) for b in range(1, length + 1, chunk)]


def pilon_name(name):
    """processRegions' newName (GenomeFile.scala:137-140)."""
    sep = "_" if "|" not in name else ("" if name.endswith("|") else "|")
    return name + sep + "pilon"


def fasta_element(header, seq):
    """writeFastaElement (GenomeFile.scala:79-82): println(">" + header), then
    sequence.sliding(80, 80) foreach println (the last group shorter, no
    group for an empty sequence)."""
    out = [">" + header + "\n"]
    for i in range(0, len(seq), 80):
        out.append(seq[i:i + 80] + "\n")
    return "".join(out)


_TARGET = re.compile(rb"\s*([^\n\r]+):([0-9]+)-([0-9]+)\s*")


def _java_split_comma(s):
    parts = s.split(",")
    if len(parts) > 1:
        while parts and parts[-1] == "":
            parts.pop()
    return parts


def parse_target_string(s, contigs):
    """parseTargetString (GenomeFile.scala:215-236) over [(name, length)]:
    [(contig index, start, stop)]; raises ValueError where Pilon throws
    (or builds a region it cannot process)."""
    index = {}
    for i, (name, _) in enumerate(contigs):
        index.setdefault(name, i)
    out = []
    for t in _java_split_comma(s):
        m = _TARGET.fullmatch(t.encode())
        if m:
            name, a, b = m.group(1).decode(), int(m.group(2)), int(m.group(3))
            if a >= 2 ** 31 or b >= 2 ** 31:
                raise ValueError("NumberFormatException: " + t)
        else:
            name = t
        if name not in index:
            raise ValueError("NoSuchElementException: " + name)
        length = contigs[index[name]][1]
        if not m:
            a, b = 1, length
        if b > length or a < 1 or b < a:
            raise ValueError("bad region " + t)
        out.append((index[name], a, b))
    return out


def parse_targets(s, contigs):
    """parseTargets (GenomeFile.scala:238-250): the file's lines if it reads
    and parses to something, else the string."""
    try:
        with open(s, "rb") as f:
            lines = f.read().decode().splitlines()
        got = [t for line in lines for t in parse_target_string(line, contigs)]
        if got:
            return got
    except (OSError, ValueError):
        pass
    return parse_target_string(s, contigs)


def process_regions_output(cfg, contigs, regions, readsets_of, element=None, set_types=None):
    """processRegions' output loop (GenomeFile.scala:135-162) over regions
    [(contig index, start, stop)] with contigs [(name, bases)]: each region
    through run_region with its element's newName and the offset of the
    element's earlier regions (:149-153); returns (.changes text, FASTA
    text).  readsets_of(i) gives region i's read batches; element (one id
    per region, None = the contig) groups consecutive regions; set_types
    are the BAM types of every region's read sets (None: frags)."""
    changes, fasta = [], []
    i = 0
    while i < len(regions):
        j = i
        key = element[i] if element is not None else regions[i][0]
        while j < len(regions) and regions[j][0] == regions[i][0] and \
                (element[j] if element is not None else regions[j][0]) == key:
            j += 1
        name, bases = contigs[regions[i][0]]
        new_name = pilon_name(name)
        offset = 0
        seq = []
        for k in range(i, j):
            _, s, e = regions[k]
            r = run_region(cfg, bases, s, e, readsets_of(k), name=name, new_name=new_name, offset=offset,
                           set_types=set_types)
            if r["status"] != 0:
                raise ValueError(f"region {k}: status {r['status']}")
            changes.append(r["changes"])
            seq.append(r["bases"].decode())
            offset += len(r["bases"]) - (e - s + 1)
        fasta.append(fasta_element(new_name, "".join(seq)))
        i = j
    return "".join(changes), "".join(fasta)
