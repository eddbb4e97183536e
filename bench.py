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
# serving lanes start a launch beside a running one only once it has run this long (microseconds): at
# 20 req/s on this workload 250 us keeps the p99 of one request at a time (2.8 ms against 4.2 ms with no
# delay) and the saturated throughput of two lanes
SERVE_OVERLAP_US = 250

# HBM bytes per launch of the dominant kernels on the plain_large workload, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs calibrated against a device copy, and vector instructions per launch from SQ_INSTS_VALU (the
# rocprofv3 summaries under profiles/, named for the kernel and workload). Count-only sequencing is bound by
# vector instruction issue, not HBM; the probe kernels' instruction counts per lane-iteration come from the
# same counter (profiles/k_probe_valu.md).
PROFILED = {"k_seq_count": dict(traffic=302.1e6, valu_insts=6.206e9),
            "k_seq_reads": dict(traffic=77.42e9),
            # batched read materialisation of the 256-job batched workload's leading 19 jobs
            "k_seq_reads_batched": dict(traffic=72.54e9),
            # wave64 VALU instructions per wave per loop iteration of the probe kernels
            "k_probe_valu": dict(philox_per_iter=43.0, int_per_iter=24.0)}

# Algorithmic bytes of one materialised read: an 8-byte header (start, source) and the 2-bit bases,
# 8 + ceil(L / 4) = 46 B at L = 150. The stored row is 8 + 4 * ceil(L / 16) = 48 B (word-aligned bases);
# the difference is counted in `traffic`, not in `achieved`.


def read_algo_bytes(L: int) -> int:
    return 8 + (L + 3) // 4


def cpu_info() -> dict:
    """The host the CPU baseline ran on: logical CPUs, those this process may
    use, the OpenMP thread setting and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return dict(cpu_count=os.cpu_count(), usable=usable, omp_num_threads=os.environ.get("OMP_NUM_THREADS"),
                model=model)


def cpu_threads(info: dict) -> int:
    """Every CPU this process may use, capped by OMP_NUM_THREADS when set
    (the host's CPU share for one GPU)."""
    n = info["usable"]
    omp = info.get("omp_num_threads")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def valu_ceiling(ctx, reps: int = 3) -> dict:
    """The vector-ALU issue ceiling measured on this GPU by the probe kernels
    (psx_probe_valu): HIP-event time of a launch of every lane running a
    Philox loop, and of eight independent integer chains per lane."""
    from process_amd import core

    out = {}
    for mode, iters in (("philox", 4096), ("int", 16384)):
        core.probe_valu(ctx, mode, iters)
        ms = min(core.probe_valu(ctx, mode, iters)[0] for _ in range(reps))
        lane_iters = core.probe_valu(ctx, mode, iters)[1]
        per_iter = PROFILED["k_probe_valu"][f"{mode}_per_iter"]
        r = dict(ms=ms, lane_iters=lane_iters, lane_iters_per_s=lane_iters / (ms / 1e3))
        if per_iter:
            r["valu_insts"] = lane_iters / 64 * per_iter
            r["valu_insts_per_s"] = r["valu_insts"] / (ms / 1e3)
        out[mode] = r
    return out


def place_leg(ctx, w, built, reps: int = 5) -> dict:
    """place_mutations on the plain-large workload, repeated: wall clock per
    call (host set-up, kernels, sorts, the table's device index) and the
    HIP-event time of the placement kernels."""
    from process_amd import core

    walls, kms = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        m = core.Mutations.place(ctx, built.genome, built.forest, w.place_seed, **w.engine)
        walls.append((time.perf_counter() - t0) * 1e3)
        kms.append(ctx.last_place_ms())
        m.close()
    ix = built.genome.context_index(w.engine.get("context_sampling", 1))
    return dict(mutations=built.muts.n, nodes=len(w.forest["cell_id"]), wall_ms=float(np.median(walls)),
                kernel_ms=float(np.median(kms)),
                value=built.muts.n / (np.median(walls) / 1e3), unit="mutations/s",
                context_index=dict(sampling=int(w.engine.get("context_sampling", 1)),
                                   sites=int(ix["kept"].sum()), build_ms=ix["build_ms"],
                                   note="built once per genome and sampling, before the timed calls"))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(w, built, recs, threads: int, target_reads: float):
    """The oracle (CPU restatement) on a bounded sample of the same workload:
    every output sample over the first chromosomes, up to target_reads reads."""
    from oracle import oracle as O

    fd = w.forest
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    seq = dict(w.seq)
    L = seq["read_size"]
    S = fd["n_samples"] + (1 if seq.get("with_normal_sample", True) else 0)
    chrs, reads = [], 0.0
    for c, ln in enumerate(w.lengths):
        chrs.append(c)
        reads += S * seq["coverage"] * int(ln) / L
        if reads >= target_reads:
            break
    seq.update(chromosomes=chrs)
    used = O.set_threads(threads)
    t0 = time.perf_counter()
    _, _, _, n = O.simulate_seq(w.lengths, of, recs, w.seq_seed, **seq)
    dt = time.perf_counter() - t0
    O.set_threads(0)
    return dict(value=n / dt, unit="reads/s", cores=used, kind="port",
                sample=f"oracle count-only simulate_seq, all {S} output samples, chromosomes 0..{chrs[-1]} of "
                       f"{len(w.lengths)}, {n} reads in {dt:.2f} s")


def cpu_reference_shaped(w, packed, recs, threads: int, target_reads: float):
    """The reference-shaped CPU restatement (each read built as base and
    quality strings, mutations and statistics in ordered maps, as RACES'
    ReadSimulator works; same reads and counts as the oracle) on a bounded
    sample of the same workload: every output sample over the first
    chromosomes, up to target_reads reads."""
    from oracle import oracle as O

    fd = w.forest
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    seq = dict(w.seq)
    L = seq["read_size"]
    S = fd["n_samples"] + (1 if seq.get("with_normal_sample", True) else 0)
    chrs, reads = [], 0.0
    for c in np.argsort(w.lengths, kind="stable"):  # shortest chromosomes first: a short bounded sample
        chrs.append(int(c))
        reads += S * seq["coverage"] * int(w.lengths[c]) / L
        if reads >= target_reads:
            break
    seq.update(chromosomes=sorted(chrs))
    t0 = time.perf_counter()
    _, _, n = O.refshape_seq(w.lengths, packed, of, recs, w.seq_seed, threads, **seq)
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="reads/s", cores=threads, kind="port",
                sample=f"reference-shaped restatement (read strings, ordered maps), all {S} output samples, "
                       f"chromosomes {sorted(chrs)} of {len(w.lengths)}, {n} reads in {dt:.2f} s")


def batched_place_leg(ctx, B, reps: int = 3) -> dict:
    """Every job's place_mutations of the batched workload: one batched call
    (the first, in the workload's construction, and again with the context's
    buffers grown) against one call per job."""
    from process_amd import core

    eng = B.engine
    t0 = time.perf_counter()
    for s, f in zip(B.specs, B.forests):
        core.Mutations.place(ctx, B.genome, f, s["place_seed"], **eng).close()
    place_one_ms = (time.perf_counter() - t0) * 1e3
    seeds = [s["place_seed"] for s in B.specs]
    warm_wall, warm_ev = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        for m in core.Mutations.place_batch(ctx, B.genome, B.forests, seeds, [eng] * len(B.specs)):
            m.close()
        warm_wall.append((time.perf_counter() - t0) * 1e3)
        warm_ev.append(ctx.last_place_ms())
    warm_wall_ms, warm_ev_ms = float(np.median(warm_wall)), float(np.median(warm_ev))
    log(f"batched: placement one call per job {place_one_ms:.1f} ms, batched again {warm_wall_ms:.1f} ms "
        f"(kernels and copies {warm_ev_ms:.1f} ms)")
    return dict(batched_wall_ms=B.place_wall_ms, batched_event_ms=B.place_ms, warm_wall_ms=warm_wall_ms,
                warm_event_ms=warm_ev_ms, one_call_per_job_wall_ms=place_one_ms,
                mutations=int(sum(m.n for m in B.muts)),
                note="every job's place_mutations in one psx_place_mutations_batch call; the first call "
                     "grows the context's buffers")


def batched_leg(ctx, n_jobs: int, reps: int, tune: bool) -> dict:
    """The batched workload: every job in one persistent hipGraph launch
    against the same jobs launched one by one (each call waits for its
    results); then the autotuner's configuration."""
    from process_amd import _abi as A
    from process_amd import core, workloads

    t = time.perf_counter()
    B = workloads.build_batched(ctx, n_jobs=n_jobs)
    leaves = [s["leaves"] for s in B.specs]
    log(f"batched: {n_jobs} jobs built in {time.perf_counter() - t:.1f} s, {B.reads:.3e} reads, "
        f"leaves {min(leaves)}..{max(leaves)}, batched placement {B.place_wall_ms:.1f} ms "
        f"(kernels and copies {B.place_ms:.1f} ms)")
    place = batched_place_leg(ctx, B)

    def timed_batch(cfg=None):
        core.simulate_seq_batch(ctx, B.jobs, cfg)  # warm-up
        walls, kms = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            n = core.simulate_seq_batch(ctx, B.jobs, cfg)
            walls.append(time.perf_counter() - t0)
            kms.append(ctx.last_kernel_ms())
        return n, float(np.median(walls)), float(np.median(kms))

    n, wall, kms = timed_batch()
    # the same launch prepared once: descriptors uploaded and graph instantiated at creation
    t0 = time.perf_counter()
    plan = core.SeqPlan(ctx, B.jobs)
    plan_setup = time.perf_counter() - t0
    plan.run()
    pw, pk = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        plan.run()
        pw.append(time.perf_counter() - t0)
        pk.append(ctx.last_kernel_ms())
    plan.close()
    plan_wall = float(np.median(pw))
    check = [k for k in range(0, n_jobs, max(1, n_jobs // 8))]
    ref = {k: B.jobs[k].results() for k in check}
    # one job at a time: each call builds descriptors, launches and synchronises
    lat = []
    t0 = time.perf_counter()
    for j in B.jobs:
        t1 = time.perf_counter()
        core.simulate_seq_device(ctx, j.genome, j.forest, j.muts, j.seed, j.alt, j.depth, j.present, **j.seq)
        lat.append(time.perf_counter() - t1)
    seq_wall = time.perf_counter() - t0
    same = all(all(np.array_equal(a, b) for a, b in zip(ref[k], B.jobs[k].results())) for k in check)
    out = dict(workload=f"{n_jobs} jobs, leaves log-uniform 1e3-1e5, coverage log-uniform 30-300x, shared 50 Mb "
                        "reference, PE 150 bp, error 1e-3, tumour + normal sample each",
               reads=n, batched=dict(value=n / wall, unit="reads/s", wall_ms=wall * 1e3, graph_ms=kms,
                                     config="default", note="descriptors built and graph captured per call"),
               plan=dict(value=n / plan_wall, unit="reads/s", wall_ms=plan_wall * 1e3,
                         graph_ms=float(np.median(pk)), setup_ms=plan_setup * 1e3, config="default",
                         note="prepared once, hipGraph replayed per run"),
               sequential=dict(value=n / seq_wall, unit="reads/s", wall_ms=seq_wall * 1e3,
                               job_p50_ms=float(np.percentile(lat, 50) * 1e3),
                               job_p99_ms=float(np.percentile(lat, 99) * 1e3)),
               batched_over_sequential=seq_wall / wall, results_identical=bool(same),
               placement=place)
    log(f"batched: {n / wall:.3e} reads/s ({wall * 1e3:.1f} ms), replayed plan {n / plan_wall:.3e} reads/s "
        f"({plan_wall * 1e3:.1f} ms), one by one {n / seq_wall:.3e} reads/s ({seq_wall * 1e3:.1f} ms), "
        f"identical {same}")
    out["plan_over_sequential"] = seq_wall / plan_wall
    reads = batched_reads_leg(ctx, B, reps)
    if reads:
        out["reads"] = reads
    if tune:
        t0 = time.perf_counter()
        cfg, best_ms = core.autotune(ctx, B.jobs)
        tune_s = time.perf_counter() - t0
        cfg.use_graph = 1
        plan = core.SeqPlan(ctx, B.jobs, cfg)
        plan.run()
        pw, pk = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            n = plan.run()
            pw.append(time.perf_counter() - t0)
            pk.append(ctx.last_kernel_ms())
        plan.close()
        wall, kms = float(np.median(pw)), float(np.median(pk))
        out["tuned"] = dict(value=n / wall, unit="reads/s", wall_ms=wall * 1e3, graph_ms=kms, tune_s=tune_s,
                            config=dict(block_threads=cfg.block_threads, lds_records=cfg.lds_records,
                                        strata_per_item=cfg.strata_per_item, blocks_per_cu=cfg.blocks_per_cu))
        out["tuned_plan_over_sequential"] = seq_wall / wall
        log(f"batched plan, tuned {out['tuned']['config']}: {n / wall:.3e} reads/s (tuning {tune_s:.1f} s)")
    for j in B.jobs:
        for a in (j.alt, j.depth, j.present):
            a.close()
    return out


def sam_leg(ctx, w, built, coverage: float = 10.0) -> dict:
    """psx_write_sam (the write_SAM = TRUE path: reads made on the GPU one
    chromosome at a time, formatted on host threads from pinned chunks) on a
    bounded sample of the workload: its shortest chromosome at a reduced
    coverage, every output sample, written to a temporary directory that is
    removed afterwards."""
    import ctypes as C
    import shutil
    import tempfile

    from process_amd import _abi as A
    from process_amd import core
    from process_amd._lib import check, lib

    c = int(np.argmin(w.lengths))
    seq = dict(w.seq, coverage=coverage, chromosomes=[c])
    S = core.num_out_samples(built.forest, **seq)
    p, keep = A.seq_params(**seq)
    names = [f"S{k}".encode() for k in range(S)]
    chrs = [f"chr{k + 1}".encode() for k in range(len(w.lengths))]
    sn = (C.c_char_p * S)(*names)
    cn = (C.c_char_p * len(chrs))(*chrs)
    out_dir = tempfile.mkdtemp(prefix="psx_sam_")
    try:
        sp = A.SamParams(out_dir.encode(), b"chr_", b"r", 0, 0, C.cast(sn, C.POINTER(C.c_char_p)),
                         C.cast(cn, C.POINTER(C.c_char_p)))
        n = C.c_uint64()
        t0 = time.perf_counter()
        check(lib().psx_write_sam(ctx.h, built.genome.h, built.forest.h, built.muts.h, C.byref(p),
                                  C.c_uint64(w.seq_seed), C.byref(sp), C.byref(n)), ctx.h)
        dt = time.perf_counter() - t0
        size = sum(os.path.getsize(os.path.join(out_dir, f)) for f in os.listdir(out_dir))
    finally:
        shutil.rmtree(out_dir, ignore_errors=True)
    del keep
    r = dict(value=n.value / dt, unit="reads/s", seconds=dt, reads=n.value, bytes=size, mb_per_s=size / dt / 1e6,
             sample=f"chromosome {c} ({int(w.lengths[c])} bases) at {coverage:g}x, {S} output samples, one SAM file",
             note="bound by host text formatting and the file system, not the GPU")
    log(f"sam: {n.value} reads, {size / 1e6:.0f} MB in {dt:.2f} s: {r['value']:.3e} reads/s, {r['mb_per_s']:.0f} MB/s")
    return r


def batched_reads_leg(ctx, B, reps: int) -> dict:
    """Read materialisation of the batched workload's leading jobs whose
    reads fit 1.6e9 rows (76.8 GB), all in one launch."""
    from process_amd import core

    sel, tot = [], 0
    for j in B.jobs:
        r = core.num_reads(j.genome, j.forest, **j.seq)
        if tot + r > 1_600_000_000:
            break
        sel.append(j)
        tot += r
    if sel:
        rw = core.read_words(**sel[0].seq)
        buf = core.DeviceArray(ctx, (tot, rw), np.uint32)
        core.simulate_seq_reads_batch(ctx, sel, buf)
        rk, rwl = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            _, first, nr = core.simulate_seq_reads_batch(ctx, sel, buf)
            rwl.append(time.perf_counter() - t0)
            rk.append(ctx.last_kernel_ms())
        buf.close()
        rkm, rwall = float(np.median(rk)), float(np.median(rwl))
        algo = read_algo_bytes(sel[0].seq["read_size"])
        ach = nr * algo / (rkm / 1e3) / 1e9
        res = dict(workload=f"read materialisation of the first {len(sel)} jobs in one launch",
                   jobs=len(sel), reads=nr, value=nr / rwall, unit="reads/s", wall_ms=rwall * 1e3,
                   kernel="k_seq_reads", kernel_ms=rkm,
                   roofline=dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                                 frac=ach / HBM_PEAK_GBS,
                                 traffic=PROFILED["k_seq_reads_batched"]["traffic"]
                                 if len(B.jobs) == 256 and len(sel) == 19 else None,
                                 algo_bytes_per_read=algo, row_bytes=4 * rw,
                                 row_write_rate_gbs=nr * 4 * rw / (rkm / 1e3) / 1e9))
        log(f"batched reads: {len(sel)} jobs, {nr} reads, kernel {rkm:.2f} ms, {ach:.0f} GB/s of read bytes "
            f"({ach / HBM_PEAK_GBS:.1%})")
        return res
    return {}


def serving_leg(ctx, rate: float, duration: float, saturated: int, overlap_us: int = 0,
                compare_overlap_us=(), forms=("continuous", "one_at_a_time")) -> dict:
    """The serving workload: Poisson arrivals of requests over device-resident
    forests (1e3-1e6 cells) through the continuous-batching server, against
    the same stream served one request at a time (forms picks either or both);
    then the same number of requests submitted at once (saturated) in both
    forms. Peak HBM is the device's whole use (free memory sampled while
    waiting on each request): resident inputs, the server's output arena and
    the context's pools."""
    import threading

    from process_amd import core, workloads

    f0, _ = ctx.mem_info()
    t = time.perf_counter()
    pool = workloads.build_serving_pool(ctx)
    f1, _ = ctx.mem_info()
    log(f"serving: {len(pool.forests)} forests ({pool.cells[0]}..{pool.cells[-1]} cells) resident in "
        f"{time.perf_counter() - t:.1f} s, {(f0 - f1) / 2**20:.0f} MiB")
    budget = 2 << 30

    def run(stream, lanes, max_batch, paced, label, overlap=0):
        srv = core.Server(ctx, budget, max_inflight=lanes, max_batch=max_batch, overlap_after_us=overlap)
        tickets = []
        t0 = time.perf_counter()

        def producer():
            for ta, k, seed, seq in stream:
                if paced:
                    dt = t0 + ta - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                tickets.append(srv.submit(pool.genome, pool.forests[k], pool.muts[k], seed, **seq))

        th = threading.Thread(target=producer)
        th.start()
        lat, queued, batch, reads = [], [], [], 0
        last = time.perf_counter()
        min_free = ctx.mem_info()[0]
        for i in range(len(stream)):
            while len(tickets) <= i:
                time.sleep(0.0005)
            min_free = min(min_free, ctx.mem_info()[0])
            _, _, _, st = srv.wait(tickets[i])
            lat.append(st["latency_ms"])
            queued.append(st["queued_ms"])
            batch.append(st["batch_size"])
            reads += st["n_reads"]
            if time.perf_counter() - last > 10:
                last = time.perf_counter()
                log(f"serving ({label}): {i + 1}/{len(stream)} requests done")
        th.join()
        wall = time.perf_counter() - t0
        tot = srv.totals()
        srv.close()
        r = dict(requests=len(stream), reads=reads, value=reads / wall, unit="reads/s", wall_s=wall,
                 p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                 queued_p99_ms=float(np.percentile(queued, 99)), mean_batch=float(np.mean(batch)),
                 launches=tot["launches"], peak_output_bytes=tot["peak_bytes"],
                 peak_device_bytes=int(ctx.mem_info()[1] - min_free))
        log(f"serving ({label}): {r['value']:.3e} reads/s, p50 {r['p50_ms']:.1f} ms, p99 {r['p99_ms']:.1f} ms, "
            f"mean batch {r['mean_batch']:.1f}")
        return r

    stream = workloads.serving_requests(pool, int(rate * duration), rate)
    # warm-up outside the timed runs: the first request of each kind loads its kernels and builds its tables
    warm = workloads.serving_requests(pool, 4 * len(pool.forests), 0.0, seed=99)
    run(warm, 2, 16, False, "warm-up", overlap_us)
    run(warm, 1, 1, False, "warm-up one at a time")
    out = dict(workload=f"Poisson arrivals at {rate:g} req/s for {duration:g} s over {len(pool.forests)} resident "
                        "forests of 1e3-1e6 cells (2-3 samples of <= 1e4 cells), shared 50 Mb reference, "
                        "coverage log-uniform 30-500x, PE 150 bp, error 1e-3, random sample subsets + normal",
               resident_input_bytes=int(f0 - f1), hbm_budget_bytes=budget,
               overlap_after_us=overlap_us)
    if "continuous" in forms:
        out["continuous"] = run(stream, 2, 16, True, "continuous", overlap_us)
    if "one_at_a_time" in forms:
        out["one_at_a_time"] = run(stream, 1, 1, True, "one at a time")
    for us in compare_overlap_us:
        out.setdefault("continuous_by_overlap_us", {})[str(us)] = run(stream, 2, 16, True,
                                                                       f"continuous, overlap after {us} us", us)
    if saturated:
        burst = workloads.serving_requests(pool, saturated, 0.0, seed=10)
        out["saturated"] = dict(requests=saturated,
                                continuous=run(burst, 2, 16, False, "saturated continuous", overlap_us),
                                one_at_a_time=run(burst, 1, 1, False, "saturated one at a time"))
    return out


def cna_leg(ctx, steps: int) -> dict:
    """The copy-number workload (count-only, purity 0.6)."""
    from process_amd import core, workloads

    t = time.perf_counter()
    w = workloads.cna_purity()
    built = workloads.build(ctx, w)
    M = built.muts.n
    ev = built.muts.cna()
    seq = w.seq
    S = core.num_out_samples(built.forest, **seq)
    alt = core.DeviceArray(ctx, (S, M), np.int32)
    dep = core.DeviceArray(ctx, (S, M), np.int32)
    pres = core.DeviceArray(ctx, (S, M), np.uint8)
    log(f"cna: {M} mutations, {len(ev)} copy-number events, built in {time.perf_counter() - t:.1f} s")
    core.simulate_seq_device(ctx, built.genome, built.forest, built.muts, w.seq_seed, alt, dep, pres, **seq)
    kms, n = [], 0
    t0 = time.perf_counter()
    for _ in range(steps):
        n += core.simulate_seq_device(ctx, built.genome, built.forest, built.muts, w.seq_seed, alt, dep, pres, **seq)
        kms.append(ctx.last_kernel_ms())
    dt = (time.perf_counter() - t0) / steps
    out = dict(workload="cna_purity: 1e5-cell two-species forest, one sample of 1e4 cells + normal, 100 Mb / 10 chr, "
                        "12 allele-specific CNA segments (1-50 Mb, states 0/1/3/4) + WGD on the driver clone, "
                        "purity 0.6, 500x PE 150 bp, error 1e-3, count-only",
               value=n / steps / dt, unit="reads/s", reads_per_step=n // steps, ms_per_step=dt * 1e3,
               kernel_ms=float(np.mean(kms)), mutations=M, cna_events=len(ev))
    log(f"cna: {n // steps} reads/step, kernel {out['kernel_ms']:.2f} ms, {out['value']:.3e} reads/s")
    for a in (alt, dep, pres):
        a.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-reads", action="store_true", help="skip the read-materialising mode")
    ap.add_argument("--cpu-reads", type=float, default=1.2e9, help="reads in the bounded CPU sample")
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched workload")
    ap.add_argument("--batched-jobs", type=int, default=256)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (plain and batched workloads)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-cna", action="store_true", help="skip the copy-number workload")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="request arrivals per second")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="length of the arrival stream")
    ap.add_argument("--serve-burst", type=int, default=200, help="requests of the saturated run (0 skips it)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run device-to-device copies of BYTES (reference for the HBM byte counters)")
    ap.add_argument("--place-only", action="store_true",
                    help="only build the plain-large workload and repeat its placement (for profiling)")
    ap.add_argument("--probe-only", action="store_true",
                    help="only run the vector-ALU probe kernels (for profiling their instruction counts)")
    ap.add_argument("--serve-overlap-us", type=int, default=SERVE_OVERLAP_US,
                    help="serving lanes start a launch beside a running one only after it ran this long")
    ap.add_argument("--serve-compare-overlap-us", type=str, default="",
                    help="comma-separated overlap settings also run on the paced stream (comparison)")
    ap.add_argument("--serving-only", action="store_true", help="only run the serving workload")
    ap.add_argument("--serve-forms", type=str, default="continuous,one_at_a_time",
                    help="forms run on the paced stream (comma-separated: continuous, one_at_a_time)")
    ap.add_argument("--cna-only", action="store_true", help="only run the copy-number workload (for profiling)")
    ap.add_argument("--no-sam", action="store_true", help="skip the SAM writer leg")
    ap.add_argument("--batched-place-only", action="store_true",
                    help="only run the batched workload's placements (--steps batched calls; for profiling)")
    ap.add_argument("--batched-reads-only", action="store_true",
                    help="only run the batched read materialisation (one launch per repetition; for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Sequencing would shard by whole (output sample, chromosome) segments, one share per GPU, each "
              "writing its own rows of the alt/depth tables.", file=sys.stderr)
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
    if args.cna_only:
        print(json.dumps(cna_leg(ctx, args.steps)), flush=True)
        return 0
    if args.serving_only:
        cmp = [int(x) for x in args.serve_compare_overlap_us.split(",") if x]
        print(json.dumps(serving_leg(ctx, args.serve_rate, args.serve_seconds, args.serve_burst,
                                     args.serve_overlap_us, cmp, args.serve_forms.split(","))), flush=True)
        return 0
    if args.probe_only:
        print(json.dumps(valu_ceiling(ctx)), flush=True)
        return 0
    if args.batched_place_only:
        B = workloads.build_batched(ctx, n_jobs=args.batched_jobs)
        print(json.dumps(batched_place_leg(ctx, B, max(1, args.steps))), flush=True)
        return 0
    if args.batched_reads_only:
        B = workloads.build_batched(ctx, n_jobs=args.batched_jobs)
        print(json.dumps(batched_reads_leg(ctx, B, max(1, args.steps))), flush=True)
        return 0
    t = time.perf_counter()
    w = workloads.plain_large(n_cells=args.cells)
    log(f"inputs: forest {len(w.forest['cell_id'])} nodes, genome {int(w.lengths.sum())} bases "
        f"({time.perf_counter() - t:.1f} s)")
    built = workloads.build(ctx, w, driver_species=1)
    M = built.muts.n
    log(f"placed {M} mutations in {built.place_ms:.2f} ms (GPU)")
    placement = place_leg(ctx, w, built)
    log(f"placement: {placement['wall_ms']:.2f} ms per call, kernels {placement['kernel_ms']:.2f} ms")
    if args.place_only:
        print(json.dumps(placement), flush=True)
        return 0
    seq = w.seq
    S = core.num_out_samples(built.forest, **seq)
    alt = core.DeviceArray(ctx, (S, M), np.int32)
    dep = core.DeviceArray(ctx, (S, M), np.int32)
    pres = core.DeviceArray(ctx, (S, M), np.uint8)
    vaf = core.DeviceArray(ctx, (S, M), np.float32)

    def step():
        # one simulate_seq over every output sample: alt, depth, presence and the fp32 VAF, all left in HBM
        n = core.simulate_seq_device(ctx, built.genome, built.forest, built.muts, w.seq_seed, alt, dep, pres, **seq)
        k_ms = ctx.last_kernel_ms()
        core.seq_vaf_f32(ctx, built.muts, alt, dep, pres, vaf)
        return n, k_ms

    tuned = None
    if not args.no_tune:
        # launch shape (block, LDS records, strata per item, blocks per CU) chosen by the autotuner on this workload
        t0 = time.perf_counter()
        job = core.SeqJob(ctx, built.genome, built.forest, built.muts, w.seq_seed, **seq)
        cfg, best_ms = core.autotune(ctx, [job])
        for a in (job.alt, job.depth, job.present):
            a.close()
        tuned = dict(block_threads=cfg.block_threads, lds_records=cfg.lds_records,
                     strata_per_item=cfg.strata_per_item, blocks_per_cu=cfg.blocks_per_cu)
        log(f"autotuned {tuned} in {time.perf_counter() - t0:.1f} s ({best_ms:.2f} ms)")

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    kms = []
    t0 = time.perf_counter()
    n_reads = 0
    for _ in range(args.steps):
        n, k_ms = step()
        n_reads += n
        kms.append(k_ms)
    ctx.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    reads_per_step = n_reads // max(1, args.steps)
    value = n_reads / elapsed
    kernel_ms = float(np.mean(kms))
    log(f"count-only: {reads_per_step} reads/step, {elapsed / args.steps * 1e3:.2f} ms/step, "
        f"kernel {kernel_ms:.2f} ms, {value:.3e} reads/s")

    # the drop-in psx_simulate_seq: the same step plus the copy of alt, depth and presence to host memory
    host_ms = []
    for _ in range(3):
        t2 = time.perf_counter()
        core.simulate_seq_arrays(ctx, built.genome, built.forest, built.muts, w.seq_seed, **seq)
        host_ms.append((time.perf_counter() - t2) * 1e3)
    host_out = dict(value=reads_per_step / (np.median(host_ms) / 1e3), unit="reads/s",
                    ms_per_step=float(np.median(host_ms)), table_bytes=9 * M * S,
                    note="psx_simulate_seq into fresh host arrays: the step plus the copy of alt, depth and presence "
                         "(pinned chunks on several host threads)")
    log(f"host outputs: {host_out['ms_per_step']:.2f} ms/step, {host_out['value']:.3e} reads/s")

    # algorithmic bytes of one count-only launch: mutation records read once
    # (16 B), alt + depth written per output sample (8 B), sample leaf lists (4 B/cell)
    leaves = int(sum(int((w.forest["sample"] == s).sum()) for s in range(w.forest["n_samples"])))
    algo_bytes = 16 * M + 8 * M * S + 4 * leaves
    achieved = algo_bytes / (kernel_ms / 1e3) / 1e9
    prof = PROFILED["k_seq_count"] if args.cells == 1_000_000 else {}
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s", frac=achieved / HBM_PEAK_GBS,
                    traffic=prof.get("traffic"), kernel="k_seq_count", kernel_ms=kernel_ms, algo_bytes=algo_bytes)
    # the bound that applies: vector instruction issue, against the probe kernels' rate on this GPU
    ceil = valu_ceiling(ctx)
    valu = dict(probe=ceil, note="count-only moves ~0.04 B per read; vector instruction issue bounds it")
    if prof.get("valu_insts") and ceil["int"].get("valu_insts_per_s"):
        rate = prof["valu_insts"] / (kernel_ms / 1e3)
        valu.update(kernel_valu_insts=prof["valu_insts"], achieved_insts_per_s=rate,
                    peak_insts_per_s=ceil["int"]["valu_insts_per_s"],
                    frac=rate / ceil["int"]["valu_insts_per_s"],
                    philox_peak_insts_per_s=ceil["philox"].get("valu_insts_per_s"),
                    frac_of_philox_rate=rate / ceil["philox"]["valu_insts_per_s"]
                    if ceil["philox"].get("valu_insts_per_s") else None)
        log(f"VALU issue: {rate:.3e} insts/s, {valu['frac']:.1%} of the integer probe's "
            f"{valu['peak_insts_per_s']:.3e}")
    roofline["valu"] = valu

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
                   "forest_nodes": len(w.forest["cell_id"]), "place_mutations_ms": placement["wall_ms"],
                   "launch": tuned or "default"},
        "roofline": roofline,
        "host_outputs": host_out,
        "placement": placement,
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
        algo = read_algo_bytes(seq["read_size"])
        ach = n_all * algo / (rk / 1e3) / 1e9
        out["reads_mode"] = dict(value=n_all / rdt, unit="reads/s", ms_per_step=rdt * 1e3, kernel="k_seq_reads",
                                 roofline=dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                                               frac=ach / HBM_PEAK_GBS,
                                               traffic=PROFILED["k_seq_reads"]["traffic"]
                                               if args.cells == 1_000_000 else None,
                                               kernel_ms=rk, algo_bytes_per_read=algo, row_bytes=4 * rw,
                                               row_write_rate_gbs=n_all * 4 * rw / (rk / 1e3) / 1e9))
        log(f"reads mode: {n_all} reads, kernel {rk:.2f} ms, {ach:.0f} GB/s of read bytes ({ach / HBM_PEAK_GBS:.1%})")
        # the write ceiling on this GPU: stores only, into the same buffer (psx_probe_store), best of 3 each
        ceil = {}
        for mode in ("fill", "chunks", "chunks_shifted"):
            core.probe_store(ctx, buf, mode)
            ms, nb = min(core.probe_store(ctx, buf, mode) for _ in range(3))
            ceil[mode] = nb / (ms / 1e3) / 1e9
        rate = n_all * 4 * rw / (rk / 1e3) / 1e9
        out["reads_mode"]["roofline"]["store_ceiling"] = dict(
            gbs=ceil, row_write_frac_of_chunks=rate / ceil["chunks"],
            note="stores only into the same buffer: grid-stride fill; the kernel's shape (items of 170 x 3 KiB "
                 "chunks, 1024-byte aligned); the same shape 16 bytes off alignment")
        log(f"store ceiling: fill {ceil['fill']:.0f} GB/s, chunks {ceil['chunks']:.0f} GB/s, shifted chunks "
            f"{ceil['chunks_shifted']:.0f} GB/s; the kernel's row writes {rate:.0f} GB/s = "
            f"{rate / ceil['chunks']:.1%} of the aligned chunks")
        buf.close()

    if not args.no_sam:
        try:
            out["sam"] = sam_leg(ctx, w, built)
        except (OSError, RuntimeError) as e:  # a full or read-only temporary file system ends only this leg
            out["sam"] = dict(error=str(e))
            log(f"sam: skipped ({e})")

    if not args.no_cna:
        out["cna_purity"] = cna_leg(ctx, 5)

    if not args.no_batched:
        out["batched"] = batched_leg(ctx, args.batched_jobs, 3, not args.no_tune)

    if not args.no_serving:
        out["serving"] = serving_leg(ctx, args.serve_rate, args.serve_seconds, args.serve_burst,
                                     args.serve_overlap_us)

    if not args.no_cpu:
        recs = built.muts.download()
        info = cpu_info()
        threads = cpu_threads(info)
        cb = cpu_baseline(w, built, recs, threads, args.cpu_reads)
        cb["host"] = info
        out["cpu_baseline"] = cb
        log(f"cpu baseline: {cb['value']:.3e} reads/s on {cb['cores']} threads")
        out["gpu_vs_cpu"] = value / cb["value"]
        # the reference's own shape of the work (RACES is single-threaded), then on every core
        packed = built.genome.download()
        rs1 = cpu_reference_shaped(w, packed, recs, 1, args.cpu_reads / 60)
        rsn = cpu_reference_shaped(w, packed, recs, threads, args.cpu_reads / 3)
        out["cpu_baseline_reference_shaped"] = dict(single_thread=rs1, all_cores=rsn,
                                                    gpu_vs_single_thread=value / rs1["value"],
                                                    gpu_vs_all_cores=value / rsn["value"])
        log(f"cpu reference-shaped: {rs1['value']:.3e} reads/s on 1 thread, {rsn['value']:.3e} on {threads}")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
