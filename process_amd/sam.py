"""SAM output of simulate_seq(write_SAM = TRUE) (src/seq_simulation.cpp:536-564,
vignettes/sequencing.Rmd:137-180, 272-309): the reads are materialised on the
GPU one chromosome at a time and written by the library's streaming writer
(psx_write_sam) into one file per chromosome, "<filename_prefix><chr>.sam",
with one read group per output sample. In update mode a chromosome whose file
exists gets the first free "<filename_prefix><chr>_<n>.sam"."""
from __future__ import annotations

import ctypes as C
import os

from . import _abi as A
from . import core
from ._lib import check, lib


def write_sam(pf, forest: core.Forest, sample_names: list[str], seq: dict, seed: int, output_dir: str,
              filename_prefix: str = "chr_", template_name_prefix: str = "r", update: bool = False,
              random_quality: bool = False, genome: core.Genome | None = None) -> int:
    """Writes the SAM files of one simulate_seq call (reads from `genome`,
    the engine's by default); returns the reads written."""
    eng = pf.engine
    ctx = eng.ctx
    g = genome if genome is not None else eng.genome()
    os.makedirs(output_dir, exist_ok=True)
    p, keep = A.seq_params(**seq)
    names = [n.encode() for n in sample_names]
    chrs = [n.encode() for n in eng.reference.names]
    sn = (C.c_char_p * max(1, len(names)))(*names)
    cn = (C.c_char_p * len(chrs))(*chrs)
    sp = A.SamParams(output_dir.encode(), filename_prefix.encode(), template_name_prefix.encode(), int(bool(update)),
                     int(bool(random_quality)), C.cast(sn, C.POINTER(C.c_char_p)), C.cast(cn, C.POINTER(C.c_char_p)))
    n = C.c_uint64()
    check(lib().psx_write_sam(ctx.h, g.h, forest.h, pf.mutations.h, C.byref(p), C.c_uint64(seed), C.byref(sp),
                              C.byref(n)), ctx.h)
    del keep
    return n.value
