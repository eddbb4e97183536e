"""SAM output of simulate_seq(write_SAM = TRUE): the reads are materialised on
the GPU (psx_simulate_seq_reads), then written one file per chromosome
("<filename_prefix><chr>.sam") with one read group per sample, as described
in vignettes/sequencing.Rmd:137-180. Quality strings are constant
(Phred round(-10 log10 e), capped at 40; 40 for the error-less sequencer):
the counts never depend on them."""
from __future__ import annotations

import math
import os

import numpy as np

from . import core

ASCII = np.frombuffer(b"ACGT", np.uint8)


def _quality(error_rate: float) -> str:
    q = 40 if error_rate <= 0 else min(40, int(round(-10 * math.log10(error_rate))))
    return chr(33 + q)


def decode_bases(words: np.ndarray, L: int) -> np.ndarray:
    """(n, ceil(L/16)) packed words -> (n, L) ASCII codes."""
    codes = ((words[:, :, None] >> (2 * np.arange(16, dtype=np.uint32))) & 3).reshape(words.shape[0], -1)[:, :L]
    return ASCII[codes]


def write_sam(pf, forest: core.Forest, sample_names: list[str], seq: dict, seed: int, output_dir: str,
              filename_prefix: str = "chr_", template_name_prefix: str = "r") -> list[str]:
    eng = pf.engine
    ctx = eng.ctx
    g = eng.genome()
    buf, n = core.simulate_seq_reads(ctx, g, forest, pf.mutations, seed, **seq)
    reads = buf.numpy()[:n]
    L = int(seq["read_size"])
    pe = int(seq.get("insert_size_mean", 0)) > 0
    rpt = 2 if pe else 1
    qual = _quality(float(seq.get("error_rate", 0.0))) * L
    chrs = seq.get("chromosomes") or list(range(len(eng.reference.names)))
    os.makedirs(output_dir, exist_ok=True)
    # reads are ordered (sample, chromosome, stratum, template, mate)
    per_chr: dict[int, list[tuple[str, np.ndarray, int]]] = {c: [] for c in chrs}
    idx, tmpl_base = 0, 0
    for name in sample_names:
        for c in sorted(set(chrs)):
            ln = int(eng.reference.lengths[c])
            if ln < L:
                continue
            T = int(math.floor(float(seq["coverage"]) * ln / (2 * L if pe else L) + 0.5))
            per_chr[c].append((name, reads[idx: idx + T * rpt], tmpl_base))
            idx += T * rpt
            tmpl_base += T
    paths = []
    for c, parts in per_chr.items():
        cname = eng.reference.names[c]
        path = os.path.join(output_dir, f"{filename_prefix}{cname}.sam")
        with open(path, "w") as fh:
            fh.write("@HD\tVN:1.6\tSO:unsorted\n")
            fh.write(f"@SQ\tSN:{cname}\tLN:{int(eng.reference.lengths[c])}\n")
            for name in sample_names:
                fh.write(f"@RG\tID:{name}\tSM:{name}\n")
            for name, rows, base in parts:
                if len(rows) == 0:
                    continue
                seqs = decode_bases(rows[:, 2:], L)
                starts = rows[:, 0].astype(np.int64)
                lines = []
                for k in range(len(rows)):
                    t = base + k // rpt
                    x = int(starts[k])
                    s = seqs[k].tobytes().decode()
                    if pe:
                        mate = k % rpt
                        other = int(starts[k + 1 if mate == 0 else k - 1])
                        flag = 99 if mate == 0 else 147
                        lo, hi = min(x, other), max(x, other) + L
                        tlen = hi - lo if mate == 0 else -(hi - lo)
                        lines.append(f"{template_name_prefix}{t}\t{flag}\t{cname}\t{x + 1}\t60\t{L}M\t=\t"
                                     f"{other + 1}\t{tlen}\t{s}\t{qual}\tRG:Z:{name}\n")
                    else:
                        lines.append(f"{template_name_prefix}{t}\t0\t{cname}\t{x + 1}\t60\t{L}M\t*\t0\t0\t{s}\t"
                                     f"{qual}\tRG:Z:{name}\n")
                fh.writelines(lines)
        paths.append(path)
    return paths
