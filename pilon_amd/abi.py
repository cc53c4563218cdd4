"""ctypes mirror of include/pilon_hip.h (the C ABI of libpilonhip.so).

The structures here are the boundary format: the same layouts a JVM host
would build through JNI/FFM.  Field names follow the header.
"""
import ctypes as C

PL_OK, PL_E_INVAL, PL_E_OOM, PL_E_DEVICE, PL_E_EDGE, PL_E_OVERFLOW, PL_E_BUSY, PL_E_REJECTED = range(8)
STATUS_NAMES = {0: "PL_OK", 1: "PL_E_INVAL", 2: "PL_E_OOM", 3: "PL_E_DEVICE", 4: "PL_E_EDGE",
                5: "PL_E_OVERFLOW", 6: "PL_E_BUSY", 7: "PL_E_REJECTED"}

PL_RF_NO_QUAL = 0x2000
PL_RF_MATE_SAME_REF = 0x4000
PL_RF_EXOTIC = 0x8000

# PileUp field order (PileUp.scala:26-39)
FIELDS = ["count_A", "count_C", "count_G", "count_T", "qual_A", "qual_C", "qual_G", "qual_T",
          "mqSum", "qSum", "physCov", "insertSize", "badPair", "deletions", "delQual",
          "insertions", "insQual", "clips"]
NFIELDS = len(FIELDS)
F = {name: i for i, name in enumerate(FIELDS)}

PL_CF_HOMO = 0x0001
PL_CF_HIGHCONF = 0x0002
PL_CF_CALLED = 0x0004
PL_CF_INDEL = 0x0008
PL_CF_HOMOINDEL = 0x0010
PL_CF_INSERTION = 0x0020
PL_CF_DELETION = 0x0040
PL_CF_SCOREPOS = 0x0080
PL_CF_DISP_SHIFT = 8
PL_CF_DISP_MASK = 0x0700
PL_CF_DELETED = 0x0800
PL_CF_WIDE = 0x1000
# Pilon's BAM kinds (--frags, --jumps, --unpaired, --nanopore, --pacbio): pl_region_add_reads_typed
BAM_TYPES = {"frags": 0, "jumps": 1, "unpaired": 2, "nanopore": 3, "pacbio": 4}

D_NONE, D_CONFIRMED, D_SNP, D_INS, D_DEL, D_AMB = range(6)


class PlConfig(C.Structure):
    _fields_ = [("flank", C.c_int32), ("min_mq", C.c_int32), ("min_qual", C.c_int32),
                ("default_qual", C.c_int32), ("min_depth", C.c_double),
                ("min_min_depth", C.c_int32), ("old_indel", C.c_int32),
                ("fix_snps", C.c_int32), ("fix_indels", C.c_int32), ("fix_amb", C.c_int32),
                ("iupac", C.c_int32), ("longread", C.c_int32), ("nonpf", C.c_int32),
                ("duplicates", C.c_int32)]


def default_config(**kw):
    """Pilon 1.24 defaults with --fix bases (Pilon.scala:52-73)."""
    c = PlConfig(flank=10, min_mq=0, min_qual=0, default_qual=10, min_depth=0.1,
                 min_min_depth=5, old_indel=0, fix_snps=1, fix_indels=1, fix_amb=0,
                 iupac=0, longread=0, nonpf=0, duplicates=0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


class PlReadSoa(C.Structure):
    _fields_ = [("n_reads", C.c_int64), ("pos", C.c_void_p), ("tlen", C.c_void_p),
                ("flag", C.c_void_p), ("mapq", C.c_void_p), ("cigar_off", C.c_void_p),
                ("cigar", C.c_void_p), ("seq_off", C.c_void_p), ("bases2", C.c_void_p),
                ("quals", C.c_void_p), ("n_exotic", C.c_int64), ("exo_read", C.c_void_p),
                ("exo_off", C.c_void_p), ("nib", C.c_void_p)]


class PlCall(C.Structure):
    _fields_ = [("base", C.c_uint8), ("alt", C.c_uint8), ("flags", C.c_uint16), ("q", C.c_int32)]


class PlRegionStats(C.Structure):
    _fields_ = [("read_count", C.c_int64), ("base_count", C.c_int64), ("coverage", C.c_int64),
                ("min_depth", C.c_int32), ("n_confirmed", C.c_int32), ("snps", C.c_int32),
                ("amb", C.c_int32), ("ins", C.c_int32), ("ins_bases", C.c_int32),
                ("dels", C.c_int32), ("del_bases", C.c_int32), ("n_unknown_ops", C.c_int32),
                ("aligned_bases", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PlRegionDesc(C.Structure):
    _fields_ = [("contig_id", C.c_int32), ("start", C.c_int32), ("stop", C.c_int32),
                ("readset_id", C.c_int32)]


class PlBatchOpts(C.Structure):
    _fields_ = [("window", C.c_int32), ("block", C.c_int32), ("persistent", C.c_int32),
                ("use_graph", C.c_int32), ("copy_out", C.c_int32)]


# numpy dtype matching pl_call (8 bytes)
def call_dtype():
    import numpy as np
    return np.dtype([("base", np.uint8), ("alt", np.uint8), ("flags", np.uint16), ("q", np.int32)])


class PlServeOpts(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("max_batch", C.c_int32), ("mem_budget", C.c_int64),
                ("max_queue", C.c_int32), ("window", C.c_int32), ("block", C.c_int32), ("copy_out", C.c_int32)]


class PlServeResult(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("tag", C.c_int64), ("status", C.c_int32), ("batch_size", C.c_int32),
                ("stats", PlRegionStats), ("est_bytes", C.c_int64), ("queue_ms", C.c_double),
                ("latency_ms", C.c_double), ("batch_device_ms", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["stats"] = self.stats.as_dict()
        return d


class PlBatchInfo(C.Structure):
    _fields_ = [("positions", C.c_int64), ("n_regions", C.c_int32), ("n_windows", C.c_int32),
                ("n_records", C.c_int64), ("n_candidates", C.c_int64), ("n_deferred", C.c_int64),
                ("window", C.c_int32), ("persistent", C.c_int32), ("block", C.c_int32), ("use_graph", C.c_int32),
                ("pipeline_ms", C.c_double), ("window_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PlTarget(C.Structure):
    """pl_target: one --targets region (contig = index into the names given)."""
    _fields_ = [("contig", C.c_int32), ("start", C.c_int32), ("stop", C.c_int32)]


SERVE_KEEP_PILEUP, SERVE_KEEP_FIX = 0x1, 0x2
CHUNK_SIZE = 10000000   # Pilon.chunkSize (Pilon.scala:52)


class PlWideCol(C.Structure):
    """pl_wide_col: a column whose Long qualSum leaves int32."""
    _fields_ = [("pos", C.c_int64), ("qual_sum", C.c_int64 * 4)]


class PlAsmOpts(C.Structure):
    """pl_asm_opts: the knobs of GapFiller.fillGap / fixBreak (Assembler.K,
    minDepth, breakRadius, and Pilon's trSafe, multiClosure, fixBreaks,
    gapMargin, minGap) and the "gaps" and "local" fixes of --fix."""
    _fields_ = [("k", C.c_int32), ("min_depth", C.c_int32), ("break_radius", C.c_int32), ("tr_safe", C.c_int32),
                ("multi_closure", C.c_int32), ("fix_breaks", C.c_int32), ("gap_margin", C.c_int32),
                ("min_gap", C.c_int32), ("fix_gaps", C.c_int32), ("fix_local", C.c_int32)]


class PlAsmFix(C.Structure):
    """pl_asm_fix: GapFiller's (start, ref, patch) and the tandem repeat noted."""
    _fields_ = [("start", C.c_int32), ("ref", C.c_void_p), ("ref_len", C.c_int64), ("patch", C.c_void_p),
                ("patch_len", C.c_int64), ("tandem_repeat", C.c_void_p)]
