"""Count-only sequencing at the full plain-large size (1e6-cell two-species
forest, three 1e4-cell samples and the normal sample, 300 Mb over 22
chromosomes, 200x paired-end 150 bp, error rate 1e-3: 1.6e9 reads), checked
through properties that do not depend on the size:

* every chromosome's counts depend on that chromosome only (its templates and
  errors are drawn from counters keyed by chromosome, sample and stratum), so
  the oracle run on the shortest chromosomes alone must give exactly their
  rows of the full GPU table;
* the table does not depend on the launch shape or on the run (a batched
  launch with another block size, LDS capacity and strata per item, and a
  second plain run, give the same bytes);
* the materialised reads of a chromosome (all 1.6e9 reads written, 76.8 GB)
  equal, row for row, the oracle's reads of a run on that chromosome alone:
  each (sample, chromosome) segment's rows start where the segments before it
  end;
* batched read materialisation of the batched workload's leading jobs up to
  1.6e9 reads in one launch: a job's rows equal its reads alone (and the
  oracle's reads of that job).
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core, workloads
from process_amd._lib import lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def large(ctx):
    say("[full] building plain-large: 1e6-cell forest, 300 Mb / 22 chromosomes")
    w = workloads.plain_large()
    built = workloads.build(ctx, w, driver_species=1)
    say(f"[full] {built.muts.n} mutations placed; simulating 200x count-only")
    a, d, p, n = core.simulate_seq_arrays(ctx, built.genome, built.forest, built.muts, w.seq_seed, **w.seq)
    say(f"[full] {n} reads")
    return w, built, (a, d, p, n)


def test_full_size_chromosomes_match_oracle(large):
    w, b, (a, d, p, n) = large
    assert n > 1_500_000_000
    lens = np.asarray(w.lengths, np.uint32)
    chosen = [int(c) for c in np.argsort(lens)[:2]]
    recs = b.muts.download()
    fd = w.forest
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    say(f"[full] oracle on chromosomes {chosen} ({int(lens[chosen].sum())} bases)")
    oa, od, op, on = O.simulate_seq(lens, of, recs, w.seq_seed, chromosomes=chosen, **w.seq)
    rows = np.isin(recs["chr"], chosen)
    assert rows.sum() > 1000 and on > 10_000_000
    S, M = a.shape
    oa, od, op = (x.reshape(S, M) for x in (oa, od, op))
    assert np.array_equal(d[:, rows], od[:, rows])
    assert np.array_equal(a[:, rows], oa[:, rows])
    assert np.array_equal(p[:, rows], op[:, rows])
    # the chosen chromosomes' rows hold reads; every other row of the oracle's run is empty
    assert d[:, rows].sum() > 0 and od[:, ~rows].sum() == 0


def test_full_size_launch_independent(ctx, large):
    w, b, (a, d, p, n) = large
    job = core.SeqJob(ctx, b.genome, b.forest, b.muts, w.seq_seed, **w.seq)
    for buf in (job.alt, job.depth, job.present):
        buf.fill(0xA5)
    say("[full] batched launch with 512 threads, 1024 LDS records, 4 strata per item")
    n2 = core.simulate_seq_batch(ctx, [job], A.LaunchConfig(512, 1024, 4, 4, 0))
    a2, d2, p2 = job.results()
    assert n2 == n
    assert np.array_equal(a2, a) and np.array_equal(d2, d) and np.array_equal(p2, p)
    for buf in (job.alt, job.depth, job.present):
        buf.close()
    say("[full] second plain run")
    a3, d3, p3, n3 = core.simulate_seq_arrays(ctx, b.genome, b.forest, b.muts, w.seq_seed, **w.seq)
    assert n3 == n and np.array_equal(a3, a) and np.array_equal(d3, d) and np.array_equal(p3, p)


def test_full_size_reads_match_oracle(ctx, large):
    w, b, _ = large
    seq = w.seq
    lens = np.asarray(w.lengths, np.uint32)
    c = int(np.argmin(lens))
    S = core.num_out_samples(b.forest, **seq)
    n_all = core.num_reads(b.genome, b.forest, **seq)
    per_chr = [core.num_reads(b.genome, b.forest, chromosomes=[k], **seq) // S for k in range(len(lens))]
    assert n_all == S * sum(per_chr)
    rw = core.read_words(**seq)
    say(f"[full] materialising {n_all} reads ({n_all * rw * 4 / 1e9:.1f} GB)")
    out, got = core.simulate_seq_reads(ctx, b.genome, b.forest, b.muts, w.seq_seed, **seq)
    assert got == n_all
    fd = w.forest
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    say(f"[full] oracle reads of chromosome {c} ({int(lens[c])} bases, {S * per_chr[c]} reads)")
    ref = O.simulate_reads(lens, b.genome.download(), of, b.muts.download(), w.seq_seed, S * per_chr[c],
                           chromosomes=[c], **seq)
    assert ref.shape == (S * per_chr[c], rw)
    per_sample = n_all // S
    for so in range(S):
        first = so * per_sample + sum(per_chr[:c])
        rows = np.empty((per_chr[c], rw), np.uint32)
        st = lib().psx_dev_download(ctx.h, rows.ctypes.data_as(C.c_void_p), C.c_void_p(out.ptr + first * rw * 4),
                                    C.c_uint64(rows.nbytes))
        assert st == 0
        exp = ref[so * per_chr[c]:(so + 1) * per_chr[c]]
        bad = np.nonzero((rows != exp).any(axis=1))[0]
        assert len(bad) == 0, (so, len(bad), bad[:3])
    out.close()


def test_plain_small_whole_table_matches_oracle(ctx):
    """The small plain configuration (5,000-cell forest, one 2,500-cell sample
    and the normal sample, 10 Mb, 80x paired-end, 1.07e7 reads): placement and
    the whole count table against the oracle, and the materialised reads."""
    say("[full] plain-small: placement and whole table against the oracle")
    w = workloads.plain_small()
    b = workloads.build(ctx, w)
    lens = np.asarray(w.lengths, np.uint32)
    packed = O.genome_synth(lens, w.genome_seed)
    assert np.array_equal(b.genome.download(), packed)
    fd = w.forest
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    om = O.place_mutations(lens, packed, of, w.place_seed, **w.engine)
    assert b.muts.download().tobytes() == om.tobytes()
    a, d, p, n = core.simulate_seq_arrays(ctx, b.genome, b.forest, b.muts, w.seq_seed, **w.seq)
    oa, od, op, on = O.simulate_seq(lens, of, om, w.seq_seed, **w.seq)
    assert n == on and n > 10_000_000
    assert np.array_equal(d, od.reshape(d.shape)) and np.array_equal(a, oa.reshape(a.shape))
    assert np.array_equal(p, op.reshape(p.shape))
    out, got = core.simulate_seq_reads(ctx, b.genome, b.forest, b.muts, w.seq_seed, **w.seq)
    rows = out.numpy()[:got]
    out.close()
    ref = O.simulate_reads(lens, packed, of, om, w.seq_seed, got, **w.seq)
    assert np.array_equal(rows, ref)


def test_batched_reads_full_size(ctx):
    """Batched read materialisation at the benchmark's size: the batched
    workload's leading jobs up to 1.6e9 reads (19 jobs, 72 GB) in one launch;
    the smallest jobs' rows equal their single-job reads, and the smallest
    job's rows equal the oracle's reads."""
    say("[full] batched reads: building the 256-job batched workload")
    B = workloads.build_batched(ctx, n_jobs=256)
    lens, _, specs = workloads.batched_jobs(seed=99, n_jobs=256)
    sel, tot = [], 0
    for j in B.jobs:
        r = core.num_reads(j.genome, j.forest, **j.seq)
        if tot + r > 1_600_000_000:
            break
        sel.append(j)
        tot += r
    say(f"[full] batched reads: {len(sel)} jobs, {tot} reads in one launch")
    out, first, n = core.simulate_seq_reads_batch(ctx, sel)
    assert n == tot == first[-1] and len(sel) > 8
    rw = core.read_words(**sel[0].seq)
    small = sorted(range(len(sel)), key=lambda k: first[k + 1] - first[k])[:3]
    for k in small:
        j = sel[k]
        cnt = first[k + 1] - first[k]
        rows = np.empty((cnt, rw), np.uint32)
        st = lib().psx_dev_download(ctx.h, rows.ctypes.data_as(C.c_void_p), C.c_void_p(out.ptr + first[k] * rw * 4),
                                    C.c_uint64(rows.nbytes))
        assert st == 0
        alone, got = core.simulate_seq_reads(ctx, j.genome, j.forest, j.muts, j.seed, **j.seq)
        assert got == cnt
        assert np.array_equal(rows, alone.numpy()[:cnt]), k
        alone.close()
        if k == small[0]:
            s = specs[k]
            say(f"[full] batched reads: oracle reads of job {k} ({s['leaves']} leaves, {cnt} reads)")
            fd = core.synth_forest(growth=[0.2], death=[0.01], intro_at=[0], n_cells=2 * s["leaves"], n_samples=1,
                                   sample_size=s["leaves"], seed=s["forest_seed"])
            of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
            ref = O.simulate_reads(np.asarray(lens, np.uint32), B.genome.download(), of, j.muts.download(), j.seed,
                                   cnt, **j.seq)
            assert np.array_equal(rows, ref)
    out.close()
