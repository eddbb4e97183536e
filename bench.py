#!/usr/bin/env python3
"""Benchmark: simulated reads/s of count-only simulate_seq on one MI355X.

Workload (plain, large): two species (growth 0.2; driver 0.3 introduced at
1,000 cells) grown to 1e6 cells, three samples of 1e4 cells plus the normal
sample, a 300 Mb synthetic reference over 22 chromosomes, SNV 1e-8 and indel
1e-9 per base per division, 200x paired-end 150 bp reads, error rate 1e-3.
Inputs are synthetic and resident in HBM before timing. A step is one
simulate_seq call over all four samples (1.6e9 reads).

Prints one JSON line. Also reports the read-materialising mode (the
write_SAM = TRUE analogue), the HIP-event duration of the dominant kernel,
its roofline position and the oracle timed on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(w, built, recs, threads: int, target_reads: float):
    """The oracle (CPU restatement) on a bounded sample of the same workload:
    sample 0 without the normal sample, the first chromosomes up to
    target_reads reads."""
    from oracle import oracle as O

    fd = w.forest
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    seq = dict(w.seq)
    L = seq["read_size"]
    chrs, reads = [], 0.0
    for c, ln in enumerate(w.lengths):
        chrs.append(c)
        reads += seq["coverage"] * int(ln) / L
        if reads >= target_reads:
            break
    seq.update(samples=[0], with_normal_sample=False, chromosomes=chrs)
    used = O.set_threads(threads)
    t0 = time.perf_counter()
    _, _, _, n = O.simulate_seq(w.lengths, of, recs, w.seq_seed, **seq)
    dt = time.perf_counter() - t0
    O.set_threads(0)
    return dict(value=n / dt, unit="reads/s", cores=used, kind="port",
                sample=f"oracle count-only simulate_seq, sample 0 only, chromosomes 0..{chrs[-1]}, "
                       f"{n} reads in {dt:.2f} s")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-reads", action="store_true", help="skip the read-materialising mode")
    ap.add_argument("--cpu-reads", type=float, default=1.5e8, help="reads in the bounded CPU sample")
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (reference for the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "see DESIGN.md for how the path would shard.", file=sys.stderr)
        return 2

    from process_amd import core, workloads

    ctx = core.Context(0)
    if args.calibrate_copy:
        import ctypes as C

        from process_amd._lib import check, lib
        src = core.DeviceArray(ctx, (args.calibrate_copy,), np.uint8).fill(1)
        dst = core.DeviceArray(ctx, (args.calibrate_copy,), np.uint8)
        for _ in range(3):
            check(lib().psx_dev_copy(ctx.h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), args.calibrate_copy), ctx.h)
        log(f"copied {args.calibrate_copy} bytes 3 times")
        return 0
    t = time.perf_counter()
    w = workloads.plain_large(n_cells=args.cells)
    log(f"inputs: forest {len(w.forest['cell_id'])} nodes, genome {int(w.lengths.sum())} bases "
        f"({time.perf_counter() - t:.1f} s)")
    built = workloads.build(ctx, w, driver_species=1)
    M = built.muts.n
    log(f"placed {M} mutations in {built.place_ms:.2f} ms (GPU)")
    seq = w.seq
    S = core.num_out_samples(built.forest, **seq)
    alt = core.DeviceArray(ctx, (S, M), np.int32)
    dep = core.DeviceArray(ctx, (S, M), np.int32)
    pres = core.DeviceArray(ctx, (S, M), np.uint8)

    def step():
        return core.simulate_seq_device(ctx, built.genome, built.forest, built.muts, w.seq_seed, alt, dep, pres, **seq)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    kms = []
    t0 = time.perf_counter()
    n_reads = 0
    for _ in range(args.steps):
        n_reads += step()
        kms.append(ctx.last_kernel_ms())
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    reads_per_step = n_reads // max(1, args.steps)
    value = n_reads / elapsed
    kernel_ms = float(np.mean(kms))
    log(f"count-only: {reads_per_step} reads/step, {elapsed / args.steps * 1e3:.2f} ms/step, "
        f"kernel {kernel_ms:.2f} ms, {value:.3e} reads/s")

    # algorithmic bytes of one count-only launch: mutation records read once
    # (16 B), alt + depth written per output sample (8 B), sample leaf lists (4 B/cell)
    leaves = int(sum(int((w.forest["sample"] == s).sum()) for s in range(w.forest["n_samples"])))
    algo_bytes = 16 * M + 8 * M * S + 4 * leaves
    achieved = algo_bytes / (kernel_ms / 1e3) / 1e9
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=None, kernel="k_seq_count", kernel_ms=kernel_ms, algo_bytes=algo_bytes)

    out = {
        "metric": "simulated reads/s on 1 MI355X (plain vs batched); p50/p99 job latency",
        "value": value,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "plain_large: 1e6-cell two-species forest, 3 samples x 1e4 cells + normal, "
                               "300 Mb / 22 chr, SNV 1e-8, indel 1e-9, 200x PE 150 bp, error 1e-3, count-only",
                   "reads_per_step": reads_per_step, "mutations": M, "samples": S,
                   "forest_nodes": len(w.forest["cell_id"]), "place_mutations_ms": built.place_ms},
        "roofline": roofline,
    }

    if not args.no_reads:
        # read-materialising mode over the same workload: 8 B header + 2-bit bases per read
        rw = core.read_words(**seq)
        n_all = core.num_reads(built.genome, built.forest, **seq)
        buf = core.DeviceArray(ctx, (n_all, rw), np.uint32)
        core.simulate_seq_reads(ctx, built.genome, built.forest, built.muts, w.seq_seed, buf, **seq)
        rms = []
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            core.simulate_seq_reads(ctx, built.genome, built.forest, built.muts, w.seq_seed, buf, **seq)
            rms.append(ctx.last_kernel_ms())
        ctx.synchronize()
        rdt = (time.perf_counter() - t0) / reps
        rk = float(np.mean(rms))
        bytes_per_read = 4 * rw
        ach = n_all * bytes_per_read / (rk / 1e3) / 1e9
        out["reads_mode"] = dict(value=n_all / rdt, unit="reads/s", ms_per_step=rdt * 1e3, kernel="k_seq_reads",
                                 roofline=dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                                               frac=ach / HBM_PEAK_GBS, traffic=None, kernel_ms=rk,
                                               bytes_per_read=bytes_per_read))
        log(f"reads mode: {n_all} reads, kernel {rk:.2f} ms, {ach:.0f} GB/s written ({ach / HBM_PEAK_GBS:.1%})")
        buf.close()

    if not args.no_cpu:
        recs = built.muts.download()
        threads = min(os.cpu_count() or 1, 16)
        cb = cpu_baseline(w, built, recs, threads, args.cpu_reads)
        out["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value']:.3e} reads/s on {cb['cores']} threads")
        out["gpu_vs_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
