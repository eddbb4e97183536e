"""Loads libpsx.so (built in-tree by process_amd.build) and declares its C ABI.

There is no fallback: if the library is missing, importing the product API
fails with an error naming the build command.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from . import _abi as A

LIB_PATH = Path(__file__).resolve().parent / "libpsx.so"

_lib = None


class PsxError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{A.STATUS.get(status, status)}: {message}")
        self.status = status
        self.message = message


def _declare(lib):
    vp = C.c_void_p
    pp = C.POINTER(C.c_void_p)
    st = C.c_int
    u32p_ = C.POINTER(C.c_uint32)
    sig = {
        "psx_abi_version": (C.c_int, []),
        "psx_ctx_create": (st, [C.c_int, pp]),
        "psx_ctx_destroy": (None, [vp]),
        "psx_last_error": (C.c_char_p, [vp]),
        "psx_ctx_synchronize": (st, [vp]),
        "psx_ctx_last_kernel_ms": (st, [vp, C.POINTER(C.c_double)]),
        "psx_ctx_last_place_ms": (st, [vp, C.POINTER(C.c_double)]),
        "psx_ctx_cancel": (st, [vp]),
        "psx_ctx_mem_info": (st, [vp, A.u64p, A.u64p]),
        "psx_ctx_trim": (st, [vp]),
        "psx_dev_alloc": (st, [vp, C.c_uint64, pp]),
        "psx_dev_free": (None, [vp, vp]),
        "psx_dev_upload": (st, [vp, vp, vp, C.c_uint64]),
        "psx_dev_download": (st, [vp, vp, vp, C.c_uint64]),
        "psx_dev_fill_u8": (st, [vp, vp, C.c_uint8, C.c_uint64]),
        "psx_dev_copy": (st, [vp, vp, vp, C.c_uint64]),
        "psx_genome_synth": (st, [vp, C.c_uint32, A.u32p, C.c_uint64, pp]),
        "psx_genome_upload": (st, [vp, C.c_uint32, A.u32p, A.u32p, pp]),
        "psx_genome_words": (C.c_uint64, [C.c_uint32, A.u32p]),
        "psx_genome_download": (st, [vp, vp, A.u32p, C.c_uint64]),
        "psx_genome_context_index": (st, [vp, vp, C.c_uint32, A.u64p, A.u64p, A.u32p, C.c_uint64,
                                          C.POINTER(C.c_double)]),
        "psx_genome_free": (None, [vp]),
        "psx_forest_upload": (st, [vp, C.POINTER(A.ForestDesc), pp]),
        "psx_forest_info": (st, [vp, A.u32p, A.u32p, A.u32p]),
        "psx_forest_flat": (st, [vp, A.i64p, A.u32p, A.u32p, A.u32p, A.u32p, A.u32p]),
        "psx_forest_free": (None, [vp]),
        "psx_place_mutations": (st, [vp, vp, vp, C.POINTER(A.EngineParams), C.c_uint64, pp]),
        "psx_place_mutations_batch": (st, [vp, vp, C.c_uint32, C.POINTER(vp), C.POINTER(C.POINTER(A.EngineParams)),
                                           A.u64p, C.POINTER(vp)]),
        "psx_mutations_upload": (st, [vp, vp, vp, C.c_uint64, C.c_uint32, pp]),
        "psx_mutations_cna": (st, [vp, vp, C.c_uint32, u32p_]),
        "psx_mutations_set_cna": (st, [vp, vp, vp, C.c_uint32]),
        "psx_mutations_count": (st, [vp, A.u64p]),
        "psx_mutations_download": (st, [vp, vp, vp, C.c_uint64]),
        "psx_mutations_free": (None, [vp]),
        "psx_mutations_save": (st, [vp, vp, C.c_char_p]),
        "psx_mutations_load": (st, [vp, C.c_char_p, pp]),
        "psx_insert_size_check": (st, [vp, C.c_uint32, C.c_uint32]),
        "psx_seq_num_samples": (C.c_uint32, [vp, C.POINTER(A.SeqParams)]),
        "psx_seq_num_reads": (st, [vp, vp, C.POINTER(A.SeqParams), A.u64p]),
        "psx_simulate_seq": (st, [vp, vp, vp, vp, C.POINTER(A.SeqParams), C.c_uint64, A.i32p, A.i32p, A.u8p,
                                  A.u64p]),
        "psx_simulate_seq_device": (st, [vp, vp, vp, vp, C.POINTER(A.SeqParams), C.c_uint64, vp, vp, vp, A.u64p]),
        "psx_simulate_seq_device_noalloc": (st, [vp, vp, vp, vp, C.POINTER(A.SeqParams), C.c_uint64, vp, vp, vp,
                                                 A.u64p]),
        "psx_seq_vaf_f32": (st, [vp, vp, C.c_uint32, vp, vp, vp, vp]),
        "psx_read_words": (C.c_uint32, [C.POINTER(A.SeqParams)]),
        "psx_seq_num_reads_table": (st, [vp, vp, vp, vp, C.POINTER(A.SeqParams), C.c_uint64, A.u64p]),
        "psx_simulate_seq_reads": (st, [vp, vp, vp, vp, C.POINTER(A.SeqParams), C.c_uint64, vp, C.c_uint64,
                                        A.u64p]),
        "psx_write_sam": (st, [vp, vp, vp, vp, C.POINTER(A.SeqParams), C.c_uint64, C.POINTER(A.SamParams), A.u64p]),
        "psx_simulate_seq_batch": (st, [vp, C.POINTER(A.SeqJob), C.c_uint32, C.POINTER(A.LaunchConfig), A.u64p]),
        "psx_simulate_seq_reads_batch": (st, [vp, C.POINTER(A.SeqJob), C.c_uint32, vp, C.c_uint64, A.u64p,
                                              A.u64p]),
        "psx_seq_plan_create": (st, [vp, C.POINTER(A.SeqJob), C.c_uint32, C.POINTER(A.LaunchConfig), pp]),
        "psx_seq_plan_run": (st, [vp, A.u64p]),
        "psx_seq_plan_free": (None, [vp]),
        "psx_autotune_seq": (st, [vp, C.POINTER(A.SeqJob), C.c_uint32, C.POINTER(A.LaunchConfig),
                                  C.POINTER(C.c_double)]),
        "psx_seq_table": (st, [vp, C.c_uint64, C.c_uint32, A.i32p, A.i32p, A.u8p, C.c_int32, A.u64p, A.u64p,
                               A.i32p, A.i32p, A.f64p]),
        "psx_server_create": (st, [vp, C.POINTER(A.ServerConfig), pp]),
        "psx_server_submit": (st, [vp, C.POINTER(A.Request), A.u64p]),
        "psx_server_wait": (st, [vp, C.c_uint64, C.c_double, C.POINTER(A.RequestStats)]),
        "psx_server_totals_get": (st, [vp, C.POINTER(A.ServerTotals)]),
        "psx_server_last_error": (C.c_char_p, [vp]),
        "psx_server_destroy": (None, [vp]),
        "psx_probe_valu": (st, [vp, C.c_int, C.c_uint32, C.POINTER(C.c_double), A.u64p]),
        "psx_probe_store": (st, [vp, vp, C.c_uint64, C.c_int, C.POINTER(C.c_double), A.u64p]),
        "psx_synth_forest_run": (st, [C.POINTER(A.SynthParams), C.c_uint64, pp]),
        "psx_synth_forest_nodes": (C.c_uint32, [vp]),
        "psx_synth_forest_get": (st, [vp, A.i64p, A.i64p, A.u32p, A.i32p, A.f64p]),
        "psx_synth_forest_free": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m process_amd.build` "
                              "(there is no CPU fallback for the accelerated path)")
        _lib = _declare(C.CDLL(str(LIB_PATH)))
        if _lib.psx_abi_version() != 1:
            raise ImportError("libpsx ABI version mismatch")
    return _lib


def check(status: int, ctx=None):
    if status != 0:
        msg = lib().psx_last_error(ctx)
        raise PsxError(status, msg.decode() if msg else "")
