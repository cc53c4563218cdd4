"""The C ABI: libpilonhip.so loads without a GPU, exports every function the
header declares, and the ctypes mirror has the C layouts."""
import ctypes as C
import os
import re
import subprocess

import pytest

from pilon_amd import abi
from pilon_amd.bam import PlBamStats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pilon_hip.h")
LIB = os.path.join(ROOT, "pilon_amd", "libpilonhip.so")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pl_[a-z_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ["pl_ctx_create", "pl_ref_upload", "pl_readset_upload", "pl_region_open", "pl_region_add_reads",
                 "pl_region_finish", "pl_region_fix", "pl_batch_run", "pl_last_error", "pl_ctx_destroy"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first (__graft_entry__.build())"
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_bam_library_exports_every_declared_symbol():
    """libpilonbam.so (the BAM reader) and include/pilon_bam.h."""
    lib = os.path.join(ROOT, "pilon_amd", "libpilonbam.so")
    assert os.path.exists(lib), "build first (__graft_entry__.build())"
    names = declared_functions(os.path.join(ROOT, "include", "pilon_bam.h"))
    assert "pl_bam_open" in names and "pl_bam_ref_reads" in names
    L = C.CDLL(lib)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_config_default_matches_pilon():
    lib = C.CDLL(LIB)
    c = abi.PlConfig()
    lib.pl_config_default(C.byref(c))
    d = abi.default_config()
    for name, _ in abi.PlConfig._fields_:
        assert getattr(c, name) == getattr(d, name), name
    assert (c.flank, c.default_qual, c.min_min_depth, c.min_depth) == (10, 10, 5, 0.1)


def test_no_device_is_an_error_not_a_fallback():
    """Without a usable GPU the context refuses to start (no CPU fallback)."""
    import torch  # noqa: F401  (only to know whether this host has a GPU)
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from pilon_amd.engine import Context, PilonError
    with pytest.raises(PilonError):
        Context()


def test_struct_layouts_match_c(tmp_path):
    exe = tmp_path / "abi_layout"
    subprocess.check_call(["gcc", "-O0", "-o", str(exe), os.path.join(ROOT, "tests", "abi_layout.c")])
    out = dict(l.rsplit(" ", 1) for l in subprocess.check_output([str(exe)]).decode().strip().split("\n"))
    sizes = {"pl_config": abi.PlConfig, "pl_read_soa": abi.PlReadSoa, "pl_call": abi.PlCall,
             "pl_region_stats": abi.PlRegionStats, "pl_region_desc": abi.PlRegionDesc,
             "pl_batch_opts": abi.PlBatchOpts, "pl_batch_info": abi.PlBatchInfo,
             "pl_serve_opts": abi.PlServeOpts, "pl_serve_result": abi.PlServeResult,
             "pl_bam_stats": PlBamStats, "pl_asm_opts": abi.PlAsmOpts, "pl_asm_fix": abi.PlAsmFix}
    for name, cls in sizes.items():
        assert int(out[name]) == C.sizeof(cls), name
    for key, val in out.items():
        if "." in key:
            sname, field = key.split(".")
            cls = sizes[sname]
            assert getattr(cls, field).offset == int(val), key


def test_asm_opts_default_matches_pilon():
    """Assembler.K 47, minDepth 5, breakRadius floor 3K, trSafe on,
    multiClosure and fixBreaks off, gapMargin 100000, minGap 10, and the
    gaps and local fixes on (--fix all)."""
    lib = C.CDLL(LIB)
    o = abi.PlAsmOpts()
    lib.pl_asm_opts_default(C.byref(o))
    assert [getattr(o, f) for f, _ in abi.PlAsmOpts._fields_] == [47, 5, 141, 1, 0, 0, 100000, 10, 1, 1]


@pytest.mark.parametrize("field,value", [("min_min_depth", -1), ("min_depth", float("nan")),
                                         ("min_depth", float("inf")), ("flank", -1), ("default_qual", 128)])
def test_ctx_create_rejects_bad_config(field, value):
    """pl_ctx_create validates the knobs before it touches a device: a
    negative minMinDepth, a non-finite minDepth, a negative flank or a
    defaultQual outside a Byte are PL_E_INVAL with a message, GPU or not."""
    lib = C.CDLL(LIB)
    lib.pl_ctx_create.restype = C.c_int
    lib.pl_last_error.restype = C.c_char_p
    c = abi.default_config(**{field: value})
    out = C.c_void_p()
    assert lib.pl_ctx_create(C.byref(c), 0, C.byref(out)) == 1
    assert lib.pl_last_error()
    assert not out.value
