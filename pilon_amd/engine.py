"""Python binding of libpilonhip.so with the shape of Pilon's Scala surface.

    ctx = Context(default_config())
    ctx.ref_upload(0, "ctg", draft_bytes)
    pr = PileUpRegion(ctx, 0, start, stop)      # new PileUpRegion(name, start, stop)
    pr.addReads(batch)                           # BamFile.process -> addRead per read
    pr.postProcess()                             # computePhysCov + GenomeRegion Pass 1
    pr(i).baseCount, pr(i).baseCall ...          # PileUp fields / BaseCall
    changes, bases = pr.fix()                    # identifyAndFixIssues + writeChanges

Every call goes through the C ABI; there is no Python or CPU fallback: if
the HIP library is missing or no GPU is usable, construction raises.
"""
import ctypes as C
import os
import time

import numpy as np

from .abi import (PlConfig, PlReadSoa, PlCall, PlRegionStats, PlRegionDesc, PlBatchOpts, PlBatchInfo, PlServeOpts,
                  PlServeResult, PlTarget, PlWideCol, NFIELDS, F, STATUS_NAMES, call_dtype, default_config, PL_CF_DISP_MASK,
                  PL_CF_DISP_SHIFT, SERVE_KEEP_PILEUP, SERVE_KEEP_FIX, CHUNK_SIZE, BAM_TYPES, PlAsmOpts, PlAsmFix)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpilonhip.so")

_lib = None


VCF_DIPLOID, VCF_QE = 0x2, 0x4

# Tracks.standardTracks (Tracks.scala:33-56): name and file of each track
TRACK_NAMES = ("Changes", "Unconfirmed", "Copy Number", "Coverage", "Bad Coverage", "Pct Bad", "Delta Coverage",
               "Dip Coverage", "Physical Coverage", "Clipped Alignments", "Weighted Qual", "Weighted MQ", "GC")
TRACK_FILES = ("Changes.wig", "Unconfirmed.wig", "CopyNumber.wig", "Coverage.wig", "BadCoverage.wig", "PctBad.wig",
               "DeltaCoverage.wig", "DipCoverage.wig", "PhysicalCoverage.wig", "ClippedAlignments.wig",
               "WeightedQual.wig", "WeightedMq.wig", "GC.wig")
NTRACKS = len(TRACK_NAMES)
WIG_TRACK_LINE = 0x1


def vcf_header(contigs, date, source, command, reference, vcfqe=False):
    """Vcf.writeHeader (Vcf.scala:28-70) for [(name, length)] contigs."""
    n = len(contigs)
    names = (C.c_char_p * max(1, n))(*[c[0].encode() for c in contigs])
    lens = (C.c_int64 * max(1, n))(*[c[1] for c in contigs])
    t, ln = C.c_void_p(), C.c_int64()
    _check(lib().pl_vcf_header(names, lens, n, VCF_QE if vcfqe else 0, date.encode(), source.encode(),
                               command.encode(), reference.encode(), C.byref(t), C.byref(ln)))
    try:
        return C.string_at(t, ln.value).decode()
    finally:
        lib().pl_free(t)


def _wig(fn, h, track, track_line):
    t, n, ms = C.c_void_p(), C.c_int64(), C.c_double()
    _check(fn(h, int(track), WIG_TRACK_LINE if track_line else 0, C.byref(t), C.byref(n), C.byref(ms)))
    try:
        return C.string_at(t, n.value).decode(), ms.value
    finally:
        lib().pl_free(t)


def _bed(fn, h, track_line):
    t, n, ms = C.c_void_p(), C.c_int64(), C.c_double()
    _check(fn(h, WIG_TRACK_LINE if track_line else 0, C.byref(t), C.byref(n), C.byref(ms)))
    try:
        return C.string_at(t, n.value).decode(), ms.value
    finally:
        lib().pl_free(t)


def _take_text(p, n, binary=False):
    try:
        b = C.string_at(p, n) if n else b""
        return b if binary else b.decode()
    finally:
        if p:
            lib().pl_free(p)


def fasta_element(header, seq):
    """GenomeFile.writeFastaElement (GenomeFile.scala:79-82): '>header' and
    the sequence in lines of 80."""
    seq = seq.encode() if isinstance(seq, str) else bytes(seq)
    t, n = C.c_void_p(), C.c_int64()
    _check(lib().pl_fasta_element(header.encode(), seq, len(seq), C.byref(t), C.byref(n)))
    return _take_text(t.value, n.value)


def pilon_name(name):
    """The polished contig name (GenomeFile.scala:137-140)."""
    t = C.c_void_p()
    _check(lib().pl_pilon_name(name.encode(), C.byref(t)))
    return _take_text(t.value, len(C.string_at(t.value)))


def contig_regions(length, chunk_size=CHUNK_SIZE):
    """GenomeFile.contigRegions (GenomeFile.scala:67-74): [(start, stop)]."""
    n = C.c_int32()
    _check(lib().pl_contig_regions(int(length), int(chunk_size), None, None, 0, C.byref(n)))
    st = np.zeros(max(1, n.value), np.int32)
    sp = np.zeros(max(1, n.value), np.int32)
    _check(lib().pl_contig_regions(int(length), int(chunk_size), st.ctypes.data, sp.ctypes.data, n.value,
                                   C.byref(n)))
    return [(int(st[k]), int(sp[k])) for k in range(n.value)]


def parse_targets(targets, contigs):
    """GenomeFile.parseTargets (GenomeFile.scala:215-250) over [(name,
    length)] contigs: [(contig index, start, stop)]."""
    m = len(contigs)
    names = (C.c_char_p * max(1, m))(*[c[0].encode() for c in contigs])
    lens = (C.c_int64 * max(1, m))(*[int(c[1]) for c in contigs])
    p, n = C.c_void_p(), C.c_int32()
    _check(lib().pl_parse_targets(targets.encode(), names, lens, m, C.byref(p), C.byref(n)))
    try:
        arr = C.cast(p, C.POINTER(PlTarget))
        return [(arr[k].contig, arr[k].start, arr[k].stop) for k in range(n.value)]
    finally:
        lib().pl_free(p)


# pl_text_sink: int32_t (*)(void* user, const char* data, int64_t n)
TEXT_SINK = C.CFUNCTYPE(C.c_int32, C.c_void_p, C.POINTER(C.c_char), C.c_int64)


class PilonError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (run __graft_entry__.build()); no CPU fallback exists")
        L = C.CDLL(LIB)
        P = C.POINTER
        sig = {
            "pl_config_default": (None, [P(PlConfig)]),
            "pl_last_error": (C.c_char_p, []),
            "pl_ctx_create": (C.c_int, [P(PlConfig), C.c_int, P(C.c_void_p)]),
            "pl_ctx_destroy": (None, [C.c_void_p]),
            "pl_ref_upload": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_char_p, C.c_int64]),
            "pl_readset_upload": (C.c_int, [C.c_void_p, C.c_int32, P(PlReadSoa), P(C.c_int32)]),
            "pl_readset_upload_typed": (C.c_int, [C.c_void_p, C.c_int32, P(PlReadSoa), C.c_int32, P(C.c_int32)]),
            "pl_readset_upload_many": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p]),
            "pl_ctx_synchronize": (C.c_int, [C.c_void_p]),
            "pl_loader_open": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, P(C.c_void_p)]),
            "pl_loader_staging": (C.c_int, [C.c_void_p, P(C.c_void_p), P(C.c_int64), P(C.c_void_p), P(C.c_int64)]),
            "pl_loader_push": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64]),
            "pl_loader_finish": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
            "pl_loader_close": (None, [C.c_void_p]),
            "pl_region_open": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, P(C.c_void_p)]),
            "pl_region_add_reads": (C.c_int, [C.c_void_p, P(PlReadSoa)]),
            "pl_region_add_reads_typed": (C.c_int, [C.c_void_p, P(PlReadSoa), C.c_int32]),
            "pl_region_use_readset": (C.c_int, [C.c_void_p, C.c_int32]),
            "pl_region_finish": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, P(PlRegionStats)]),
            "pl_region_fix": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64, P(C.c_void_p), P(C.c_int64),
                                        P(C.c_void_p), P(C.c_int64)]),
            "pl_region_close": (None, [C.c_void_p]),
            "pl_free": (None, [C.c_void_p]),
            "pl_batch_run": (C.c_int, [C.c_void_p, P(PlRegionDesc), C.c_int32, P(PlBatchOpts), C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
            "pl_batch_repeat": (C.c_int, [C.c_void_p, C.c_int32]),
            "pl_last_timing": (C.c_int, [C.c_void_p, P(C.c_double), P(C.c_double)]),
            "pl_last_batch_info": (C.c_int, [C.c_void_p, P(PlBatchInfo)]),
            "pl_serve_open": (C.c_int, [C.c_void_p, P(PlServeOpts), P(C.c_void_p)]),
            "pl_serve_submit": (C.c_int, [C.c_void_p, P(PlRegionDesc), C.c_int64, P(C.c_int64)]),
            "pl_serve_poll": (C.c_int, [C.c_void_p, P(PlServeResult), C.c_int32, C.c_int32, P(C.c_int32)]),
            "pl_serve_fetch": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
            "pl_serve_drain": (C.c_int, [C.c_void_p]),
            "pl_serve_close": (None, [C.c_void_p]),
            "pl_batch_vcf": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_void_p), P(C.c_int64), P(C.c_double)]),
            "pl_region_vcf": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_void_p), P(C.c_int64), P(C.c_double)]),
            "pl_batch_tracks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_double)]),
            "pl_region_tracks": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_double)]),
            "pl_batch_wig": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, P(C.c_void_p), P(C.c_int64),
                                       P(C.c_double)]),
            "pl_region_wig": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, P(C.c_void_p), P(C.c_int64),
                                        P(C.c_double)]),
            "pl_kmer_graph": (C.c_int, [C.c_void_p, P(PlReadSoa), C.c_int32, C.c_int32, P(C.c_void_p),
                                        P(C.c_void_p), P(C.c_void_p), P(C.c_int64), P(C.c_double)]),
            "pl_batch_bed": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_void_p), P(C.c_int64), P(C.c_double)]),
            "pl_region_bed": (C.c_int, [C.c_void_p, C.c_uint32, P(C.c_void_p), P(C.c_int64), P(C.c_double)]),
            "pl_vcf_header": (C.c_int, [P(C.c_char_p), P(C.c_int64), C.c_int32, C.c_uint32, C.c_char_p, C.c_char_p,
                                        C.c_char_p, C.c_char_p, P(C.c_void_p), P(C.c_int64)]),
            "pl_batch_fix": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, P(C.c_void_p), P(C.c_int64),
                                       P(C.c_void_p), P(C.c_int64)]),
            "pl_fasta_element": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int64, P(C.c_void_p), P(C.c_int64)]),
            "pl_pilon_name": (C.c_int, [C.c_char_p, P(C.c_void_p)]),
            "pl_contig_regions": (C.c_int, [C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, P(C.c_int32)]),
            "pl_parse_targets": (C.c_int, [C.c_char_p, P(C.c_char_p), P(C.c_int64), C.c_int32, P(C.c_void_p),
                                           P(C.c_int32)]),
            "pl_serve_fetch_fix": (C.c_int, [C.c_void_p, C.c_int64, C.c_char_p, C.c_int64, P(C.c_void_p),
                                             P(C.c_int64), P(C.c_void_p), P(C.c_int64)]),
            "pl_last_wide_columns": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, P(C.c_int64)]),
            "pl_batch_run_noalloc": (C.c_int, [C.c_void_p, P(PlRegionDesc), C.c_int32, P(PlBatchOpts), C.c_void_p,
                                               C.c_void_p, C.c_void_p]),
            "pl_batch_vcf_stream": (C.c_int, [C.c_void_p, C.c_uint32, TEXT_SINK, C.c_void_p, P(C.c_int64),
                                              P(C.c_double)]),
            "pl_asm_opts_default": (None, [P(PlAsmOpts)]),
            "pl_region_reassemble": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                               P(PlAsmOpts), P(C.c_int32)]),
            "pl_region_reassembly_log": (C.c_int, [C.c_void_p, P(C.c_void_p), P(C.c_int64)]),
            "pl_assemble_break": (C.c_int, [C.c_void_p, P(PlReadSoa), C.c_char_p, C.c_char_p, C.c_int64, C.c_int32,
                                            C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(PlAsmOpts), P(PlAsmFix)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise PilonError(st, lib().pl_last_error().decode(errors="replace"))


class Context:
    """pl_ctx: one GPU, the resident contigs and read sets."""

    def __init__(self, config=None, device=0):
        self.config = config if config is not None else default_config()
        h = C.c_void_p()
        _check(lib().pl_ctx_create(C.byref(self.config), device, C.byref(h)))
        self.h = h
        self.contigs = {}

    def close(self):
        if getattr(self, "h", None):
            lib().pl_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def ref_upload(self, contig_id, name, bases):
        bases = bases.encode("ascii") if isinstance(bases, str) else bytes(bases)
        _check(lib().pl_ref_upload(self.h, contig_id, name.encode(), bases, len(bases)))
        self.contigs[contig_id] = (name, bases)

    def readset_upload(self, contig_id, reads, bam_type="frags"):
        """reads: a ReadBatch or a PlReadSoa of one BAM of bam_type ("frags",
        "unpaired", "nanopore" or "pacbio"); returns the resident read-set id."""
        soa = reads.soa() if hasattr(reads, "soa") else reads
        rid = C.c_int32()
        _check(lib().pl_readset_upload_typed(self.h, contig_id, C.byref(soa), BAM_TYPES[bam_type], C.byref(rid)))
        return rid.value

    def synchronize(self):
        """Wait until the device has finished all work queued on it."""
        _check(lib().pl_ctx_synchronize(self.h))

    def batch_repeat(self, n_passes):
        """n_passes more passes of the last batch (run with use_graph=True),
        replayed back to back with one host wait at the end
        (pl_batch_repeat)."""
        _check(lib().pl_batch_repeat(self.h, int(n_passes)))

    def readset_upload_many(self, items):
        """items: (contig_id, reads, bam_type) tuples; every set uploaded in
        one pass (pl_readset_upload_many).  Returns their read-set ids."""
        items = list(items)
        n = len(items)
        if n == 0:
            return []
        soas = [r.soa() if hasattr(r, "soa") else r for _, r, _ in items]
        arr = (C.POINTER(PlReadSoa) * n)(*[C.pointer(x) for x in soas])
        cids = (C.c_int32 * n)(*[int(c) for c, _, _ in items])
        types = (C.c_int32 * n)(*[BAM_TYPES[t] for _, _, t in items])
        ids = (C.c_int32 * n)()
        _check(lib().pl_readset_upload_many(self.h, n, cids, arr, types, ids))
        return list(ids)

    def readsets_from_bam(self, path, contig_ids, bam_type="frags", threads=0, slice_bytes=64 << 20):
        """Every reference of a BAM as a resident read set, its records decoded
        on the device (pl_loader_*): the host inflates the BGZF blocks
        (libpilonbam's stream) into the loader's pinned staging buffers.
        contig_ids: the context contig of each BAM reference (-1: skip).
        Returns the read-set ids (-1 for a skipped reference)."""
        from .bam import BamStream
        contig_ids = list(contig_ids)
        ids = (C.c_int32 * max(1, len(contig_ids)))()
        t = [time.perf_counter()]
        with BamStream(path, threads=threads) as s:
            if len(contig_ids) != len(s.names):
                raise ValueError("%d contig ids for %d BAM references" % (len(contig_ids), len(s.names)))
            ld = C.c_void_p()
            t.append(time.perf_counter())
            _check(lib().pl_loader_open(self.h, max(s.total_bytes, 0), slice_bytes, C.byref(ld)))
            t.append(time.perf_counter())
            try:
                buf, offs, cap, rcap = C.c_void_p(), C.c_void_p(), C.c_int64(), C.c_int64()
                while True:
                    _check(lib().pl_loader_staging(ld, C.byref(buf), C.byref(cap), C.byref(offs), C.byref(rcap)))
                    n, r = s.next_into(buf, cap.value, offs, rcap.value)
                    if n == 0:
                        break
                    _check(lib().pl_loader_push(ld, n, r))
                t.append(time.perf_counter())
                cids = (C.c_int32 * max(1, len(contig_ids)))(*contig_ids)
                _check(lib().pl_loader_finish(ld, len(contig_ids), cids, BAM_TYPES[bam_type], ids))
                t.append(time.perf_counter())
            finally:
                lib().pl_loader_close(ld)
            self.last_bam_stats = s.stats
        t.append(time.perf_counter())
        self.last_load_s = dict(zip(["open", "loader_open", "slices", "finish", "close"],
                                    [b - a for a, b in zip(t, t[1:])]))
        return list(ids)[:len(contig_ids)]

    def batch_descs(self, regions):
        n = len(regions)
        return (PlRegionDesc * n)(*[PlRegionDesc(c, s, e, r) for c, s, e, r in regions])

    def batch_run(self, regions, window=0, persistent=0, copy_out=True, want_stats=True, descs=None, block=0,
                  use_graph=False):
        """regions: list of (contig_id, start, stop, readset_id).  Returns
        (pile [NFIELDS, G], calls [G], stats list) — arrays are None when
        copy_out is False (results stay on the device), stats is None when
        want_stats is False.  window/block 0 = autotuned (cached per batch
        shape); use_graph replays the captured device pipeline when the batch
        repeats."""
        n = len(regions)
        if descs is None:
            descs = self.batch_descs(regions)
        opts = PlBatchOpts(window=window, block=block, persistent=persistent, use_graph=int(bool(use_graph)),
                           copy_out=int(copy_out))
        stats = (PlRegionStats * n)() if want_stats else None
        pile = calls = None
        if copy_out:
            G = sum(e - s + 1 for _, s, e, _ in regions)
            pile = np.zeros((NFIELDS, G), np.int32)
            calls = np.zeros(G, call_dtype())
        _check(lib().pl_batch_run(self.h, descs, n, C.byref(opts),
                                  pile.ctypes.data if copy_out else None,
                                  calls.ctypes.data if copy_out else None, stats if want_stats else None))
        return pile, calls, ([stats[i].as_dict() for i in range(n)] if want_stats else None)

    def batch_vcf(self, diploid=False, vcfqe=False, decode=True):
        """GenomeRegion.writeVcf for every region of the last batch, in batch
        order (Vcf.writeRecord lines, formatted on the device); returns
        (text, device_ms), or (bytes of text, lines, device_ms) when decode
        is False."""
        flags = (VCF_DIPLOID if diploid else 0) | (VCF_QE if vcfqe else 0)
        t, n, ms = C.c_void_p(), C.c_int64(), C.c_double()
        t0 = time.perf_counter()
        _check(lib().pl_batch_vcf(self.h, flags, C.byref(t), C.byref(n), C.byref(ms)))
        self.last_vcf_wall_ms = (time.perf_counter() - t0) * 1000.0
        try:
            if not decode:
                buf = np.ctypeslib.as_array(C.cast(t, C.POINTER(C.c_uint8)), shape=(n.value,))
                return n.value, int(np.count_nonzero(buf == 10)), ms.value
            text = C.string_at(t, n.value).decode()
        finally:
            lib().pl_free(t)
        return text, ms.value

    def batch_fix(self, element=None, want_changes=True, want_fasta=True):
        """GenomeFile.processRegions' output loop over the last batch: (the
        .changes text, the FASTA text).  element: one id per region (None: the
        contig); consecutive regions sharing an id form one FASTA element."""
        el = None
        if element is not None:
            el = np.ascontiguousarray(element, dtype=np.int32)
        t, tl, f, fl = C.c_void_p(), C.c_int64(), C.c_void_p(), C.c_int64()
        _check(lib().pl_batch_fix(self.h, el.ctypes.data if el is not None else None, len(el) if el is not None else 0,
                                  C.byref(t) if want_changes else None, C.byref(tl),
                                  C.byref(f) if want_fasta else None, C.byref(fl)))
        changes = _take_text(t.value, tl.value) if want_changes else None
        fasta = _take_text(f.value, fl.value) if want_fasta else None
        return changes, fasta

    def batch_vcf_stream(self, write, diploid=False, vcfqe=False, raw=False):
        """batch_vcf's text handed to write(bytes) a piece at a time (a
        false return stops it); returns (bytes handed over, device_ms).  With
        raw, write(address, length) gets the library's staging memory itself,
        valid during the call."""
        flags = (VCF_DIPLOID if diploid else 0) | (VCF_QE if vcfqe else 0)

        def sink(_user, data, n):
            r = write(data, n) if raw else write(C.string_at(data, n))
            return 0 if r is not False else 1

        cb = TEXT_SINK(sink)
        n, ms = C.c_int64(), C.c_double()
        _check(lib().pl_batch_vcf_stream(self.h, flags, cb, None, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def batch_tracks(self, copy=True):
        """Tracks.standardTracks values of the last batch: ([NTRACKS, G] int32
        in batch order, or None when copy is False; device_ms)."""
        ms = C.c_double()
        vals = None
        if copy:
            vals = np.empty((NTRACKS, self.batch_info()["positions"]), np.int32)
        _check(lib().pl_batch_tracks(self.h, vals.ctypes.data if copy else None, vals.size if copy else 0,
                                     C.byref(ms)))
        return vals, ms.value

    def batch_wig(self, track, track_line=True):
        """Tracks.makeTrack text of one track over the last batch's regions:
        (text, device_ms)."""
        return _wig(lib().pl_batch_wig, self.h, track, track_line)

    def batch_bed(self, track_line=True):
        """Tracks.makeBedTrack (Pilon.bed) lines of the last batch's regions:
        (text, device_ms)."""
        return _bed(lib().pl_batch_bed, self.h, track_line)

    def kmer_graph(self, reads, K=47, min_depth=5):
        """Assembler.addReads + buildGraph on the device: ([(kmer, kGraph base,
        altGraph base or None)] in byte order of the K-mers, device_ms)."""
        soa = reads.soa() if hasattr(reads, "soa") else reads
        k, nx, al = C.c_void_p(), C.c_void_p(), C.c_void_p()
        n, ms = C.c_int64(), C.c_double()
        _check(lib().pl_kmer_graph(self.h, C.byref(soa), K, min_depth, C.byref(k), C.byref(nx), C.byref(al),
                                   C.byref(n), C.byref(ms)))
        try:
            m = n.value
            if not m:
                return [], ms.value
            ks = C.string_at(k, K * m)
            ns = C.string_at(nx, m)
            als = C.string_at(al, m)
            return [(ks[K * i:K * i + K].decode(), chr(ns[i]), chr(als[i]) if als[i] else None)
                    for i in range(m)], ms.value
        finally:
            for p in (k, nx, al):
                if p.value:
                    lib().pl_free(p)

    def assemble_break(self, reads, bases, contig, start, stop, break_start, break_stop, is_gap, **opts):
        """GapFiller.fillGap (is_gap) or fixBreak (GapFiller.scala:39-98) of
        [break_start, break_stop] in region [start, stop], given the reads
        recruited for it: bases are the region's current bases, contig the
        contig as read.  Keyword options are pl_asm_opts fields (k,
        min_depth, break_radius, tr_safe, multi_closure, fix_breaks,
        gap_margin, min_gap).  Returns ((fix start, ref, patch), tandem
        repeat); (0, "", "") when there is no fix.  Raises PilonError (EDGE)
        where Pilon 1.24 throws."""
        soa = reads.soa() if hasattr(reads, "soa") else reads
        o = PlAsmOpts()
        lib().pl_asm_opts_default(C.byref(o))
        for key, v in opts.items():
            if key not in dict(PlAsmOpts._fields_):
                raise TypeError("unknown assembly option %r" % key)
            if v is not None:
                setattr(o, key, int(v))
        b = bases.encode() if isinstance(bases, str) else bytes(bases)
        g = contig.encode() if isinstance(contig, str) else bytes(contig)
        if len(b) != stop + 1 - start:
            raise ValueError("bases hold %d bytes for region %d-%d" % (len(b), start, stop))
        out = PlAsmFix()
        _check(lib().pl_assemble_break(self.h, C.byref(soa), b, g, len(g), start, stop, break_start, break_stop,
                                       int(bool(is_gap)), C.byref(o), C.byref(out)))
        try:
            ref = C.string_at(out.ref, out.ref_len).decode()
            patch = C.string_at(out.patch, out.patch_len).decode()
            loop = C.string_at(out.tandem_repeat).decode()
        finally:
            for p in (out.ref, out.patch, out.tandem_repeat):
                if p:
                    lib().pl_free(p)
        return (out.start, ref, patch), loop

    def wide_columns(self):
        """Columns of the last batch whose Long qualSum leaves int32:
        {batch position: (qualSum A, C, G, T)} (BaseSum.scala:24)."""
        n = C.c_int64()
        _check(lib().pl_last_wide_columns(self.h, None, 0, C.byref(n)))
        if not n.value:
            return {}
        buf = (PlWideCol * n.value)()
        _check(lib().pl_last_wide_columns(self.h, buf, n.value, C.byref(n)))
        return {int(buf[k].pos): tuple(int(x) for x in buf[k].qual_sum) for k in range(n.value)}

    def batch_info(self):
        info = PlBatchInfo()
        _check(lib().pl_last_batch_info(self.h, C.byref(info)))
        return info.as_dict()

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        _check(lib().pl_last_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value


class Server:
    """pl_server: the serving form.  Requests (one region each) are admitted
    by estimated device memory and run with continuous batching on `lanes`
    concurrent batch pipelines over the context's resident data."""

    def __init__(self, ctx, lanes=2, max_batch=256, mem_budget=0, max_queue=0, window=0, block=0, copy_out=False,
                 keep_fix=False):
        self.ctx = ctx
        self.copy_out = bool(copy_out)
        keep = (SERVE_KEEP_PILEUP if copy_out else 0) | (SERVE_KEEP_FIX if keep_fix else 0)
        opts = PlServeOpts(lanes=lanes, max_batch=max_batch, mem_budget=int(mem_budget), max_queue=max_queue,
                           window=window, block=block, copy_out=keep)
        h = C.c_void_p()
        _check(lib().pl_serve_open(ctx.h, C.byref(opts), C.byref(h)))
        self.h = h
        self.sizes = {}

    def submit(self, contig_id, start, stop, readset_id, tag=0):
        """Queues one region; returns its ticket.  PilonError with status
        PL_E_REJECTED (alone over the budget) or PL_E_BUSY (queue full)."""
        t = C.c_int64()
        _check(lib().pl_serve_submit(self.h, C.byref(PlRegionDesc(contig_id, start, stop, readset_id)), int(tag),
                                     C.byref(t)))
        self.sizes[t.value] = stop - start + 1
        return t.value

    def poll(self, max_n=4096, wait_ms=0):
        buf = (PlServeResult * max_n)()
        n = C.c_int32()
        _check(lib().pl_serve_poll(self.h, buf, max_n, int(wait_ms), C.byref(n)))
        return [buf[i].as_dict() for i in range(n.value)]

    def fetch_fix(self, ticket, new_name=None, offset=0):
        """(.changes text, fixed bases) of a polled request of a keep_fix server."""
        t, tl, b, bl = C.c_void_p(), C.c_int64(), C.c_void_p(), C.c_int64()
        _check(lib().pl_serve_fetch_fix(self.h, int(ticket), new_name.encode() if new_name else None, int(offset),
                                        C.byref(t), C.byref(tl), C.byref(b), C.byref(bl)))
        return _take_text(t.value, tl.value), _take_text(b.value, bl.value, binary=True)

    def fetch(self, ticket):
        G = self.sizes[ticket]
        pile = np.zeros((NFIELDS, G), np.int32)
        calls = np.zeros(G, call_dtype())
        _check(lib().pl_serve_fetch(self.h, int(ticket), pile.ctypes.data, calls.ctypes.data))
        return pile, calls

    def drain(self):
        _check(lib().pl_serve_drain(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().pl_serve_close(self.h)
            self.h = None

    def __del__(self):
        self.close()


class BaseSum:
    """A view of one BaseSum (BaseSum.scala:23-74)."""

    def __init__(self, sums):
        self.sums = [int(x) for x in sums]

    def sum(self):
        return sum(self.sums)

    def order(self):
        return sorted(range(4), key=lambda i: -self.sums[i])  # stable: ties keep A<C<G<T

    def maxBase(self):
        o = self.order()
        return o[0], self.sums[o[0]], self.sum()


class PileUp:
    """Read-only view of one position (PileUp.scala:25-41) plus its BaseCall."""

    def __init__(self, pile_col, call):
        f = pile_col
        self.baseCount = BaseSum(f[F["count_A"]:F["count_T"] + 1])
        self.qualSum = BaseSum(f[F["qual_A"]:F["qual_T"] + 1])
        for name in ("mqSum", "qSum", "physCov", "insertSize", "badPair", "deletions", "delQual",
                     "insertions", "insQual", "clips"):
            setattr(self, name, int(f[F[name]]))
        self.call = call

    def count(self):
        return self.baseCount.sum()

    def depth(self):
        return self.count() + self.deletions

    @property
    def baseCall(self):
        c = self.call
        return {"base": "ACGTN"[c["base"]], "altBase": "ACGT"[c["alt"]], "q": int(c["q"]),
                "flags": int(c["flags"]),
                "disposition": (int(c["flags"]) & PL_CF_DISP_MASK) >> PL_CF_DISP_SHIFT}


class PileUpRegion:
    """PileUpRegion (PileUpRegion.scala:26) + the GenomeRegion Pass-1 output."""

    def __init__(self, ctx, contig_id, start, stop):
        self.ctx = ctx
        self.contig_id = contig_id
        self.start, self.stop = start, stop
        self.size = stop + 1 - start
        h = C.c_void_p()
        _check(lib().pl_region_open(ctx.h, contig_id, start, stop, C.byref(h)))
        self.h = h
        self.pile = None
        self.calls = None
        self.stats = None

    def close(self):
        if getattr(self, "h", None):
            lib().pl_region_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def addReads(self, reads, bam_type="frags"):
        """BamFile.process for one BAM's worth of reads (query, validateRead,
        addRead).  bam_type is the BAM's kind: "frags", "jumps" (left out of
        fragment coverage, GenomeRegion.scala:287-300), "unpaired", or the
        long-read "nanopore" / "pacbio" (PileUpRegion.scala:120-190)."""
        soa = reads.soa() if hasattr(reads, "soa") else reads
        _check(lib().pl_region_add_reads_typed(self.h, C.byref(soa), BAM_TYPES[bam_type]))

    def useReadSet(self, readset_id):
        _check(lib().pl_region_use_readset(self.h, readset_id))

    def postProcess(self):
        pile = np.zeros((NFIELDS, self.size), np.int32)
        calls = np.zeros(self.size, call_dtype())
        st = PlRegionStats()
        _check(lib().pl_region_finish(self.h, pile.ctypes.data, calls.ctypes.data, C.byref(st)))
        self.pile, self.calls, self.stats = pile, calls, st.as_dict()
        return self

    @property
    def readCount(self):
        return self.stats["read_count"]

    @property
    def baseCount(self):
        return self.stats["base_count"]

    @property
    def coverage(self):
        return self.stats["coverage"]

    def __call__(self, i):
        """apply(i): the PileUp at region index i (PileUpRegion.scala:233)."""
        return PileUp(self.pile[:, i], self.calls[i])

    def vcf(self, diploid=False, vcfqe=False):
        """GenomeRegion.writeVcf for this finished region: (text, device_ms)."""
        flags = (VCF_DIPLOID if diploid else 0) | (VCF_QE if vcfqe else 0)
        t, n, ms = C.c_void_p(), C.c_int64(), C.c_double()
        _check(lib().pl_region_vcf(self.h, flags, C.byref(t), C.byref(n), C.byref(ms)))
        try:
            return C.string_at(t, n.value).decode(), ms.value
        finally:
            lib().pl_free(t)

    def tracks(self):
        """Tracks.standardTracks values of this finished region: ([NTRACKS,
        size] int32, device_ms)."""
        vals = np.empty((NTRACKS, self.stop - self.start + 1), np.int32)
        ms = C.c_double()
        _check(lib().pl_region_tracks(self.h, vals.ctypes.data, vals.size, C.byref(ms)))
        return vals, ms.value

    def wig(self, track, track_line=True):
        """Tracks.makeTrack text of one track over this region: (text, device_ms)."""
        return _wig(lib().pl_region_wig, self.h, track, track_line)

    def bed(self, track_line=True):
        """Tracks.makeBedTrack (Pilon.bed) lines of this region: (text, device_ms)."""
        return _bed(lib().pl_region_bed, self.h, track_line)

    def reassemble(self, sets, break_radius=None, **opts):
        """identifyAndFixIssues' gap filling and break fixing for this finished
        region (GenomeRegion.scala:382-410, as with --nostrays).  sets: the
        region's read sets as (reads, bam_type, maxInsertSize) tuples, from
        which each break's reads are recruited; break_radius is
        GapFiller.breakRadius (3K by default); opts are other pl_asm_opts
        fields.  Returns (number of fixes kept, the logFix lines); fix() and
        bed() then include them."""
        o = PlAsmOpts()
        lib().pl_asm_opts_default(C.byref(o))
        if break_radius is not None:
            o.break_radius = int(break_radius)
        for key, v in opts.items():
            if key not in dict(PlAsmOpts._fields_):
                raise TypeError("unknown assembly option %r" % key)
            setattr(o, key, int(v))
        soas = [r.soa() if hasattr(r, "soa") else r for r, _, _ in sets]
        n = len(soas)
        arr = (C.POINTER(PlReadSoa) * max(1, n))(*[C.pointer(x) for x in soas])
        types = (C.c_int32 * max(1, n))(*[BAM_TYPES[t] for _, t, _ in sets])
        ins = (C.c_int32 * max(1, n))(*[int(m) for _, _, m in sets])
        nf = C.c_int32()
        _check(lib().pl_region_reassemble(self.h, arr, types, ins, n, C.byref(o), C.byref(nf)))
        t, tl = C.c_void_p(), C.c_int64()
        _check(lib().pl_region_reassembly_log(self.h, C.byref(t), C.byref(tl)))
        try:
            return nf.value, C.string_at(t, tl.value).decode()
        finally:
            lib().pl_free(t)

    def fix(self, new_name=None, offset=0):
        """identifyAndFixIssues + writeChanges: returns (changes text, fixed bases)."""
        t, tl, b, bl = C.c_void_p(), C.c_int64(), C.c_void_p(), C.c_int64()
        _check(lib().pl_region_fix(self.h, new_name.encode() if new_name else None, offset,
                                   C.byref(t), C.byref(tl), C.byref(b), C.byref(bl)))
        text = C.string_at(t, tl.value).decode()
        bases = C.string_at(b, bl.value)
        lib().pl_free(t)
        lib().pl_free(b)
        return text, bases
