"""Bit-exact parity of the gfx950 path with the CPU oracle (same seeds, same
inputs): reference synthesis, forest flattening, mutation placement and
count-only sequencing, including the LDS-overflow path, replayed tables,
batched launches and read materialisation."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import cases
from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core
from process_amd._lib import PsxError

pytestmark = pytest.mark.gpu

LENS = np.array([40_000, 23_457], np.uint32)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class World:
    pass


def build_world(ctx, lens, genome_seed, forest_kw, engine_kw, place_seed):
    w = World()
    w.lens = np.asarray(lens, np.uint32)
    w.packed = O.genome_synth(w.lens, genome_seed)
    w.g = core.Genome(ctx, w.lens, seed=genome_seed)
    fd = cases.forest(**forest_kw)
    w.fd = fd
    w.of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    w.pf = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    w.eng = cases.engine(w.packed, w.lens, **engine_kw)
    w.om = O.place_mutations(w.lens, w.packed, w.of, place_seed, **w.eng)
    w.pm = core.Mutations.place(ctx, w.g, w.pf, place_seed, **w.eng)
    return w


@pytest.fixture(scope="module")
def world(ctx):
    say("[parity] building world: genome 40000+23457, forest 400 cells, place seed 12345")
    return build_world(ctx, LENS, 77, dict(n_cells=400, n_samples=2, sample_size=80, seed=3), {}, 12345)


def test_genome_bit_exact(world):
    assert np.array_equal(world.g.download(), world.packed)


def test_forest_flat_matches(world):
    fl = world.pf.flat()
    of = world.of
    assert np.array_equal(fl["cell_id"], of.cell_id)
    assert np.array_equal(fl["parent"], of.parent)
    assert np.array_equal(fl["leaf_lo"], of.leaf_lo)
    assert np.array_equal(fl["leaf_hi"], of.leaf_hi)
    assert np.array_equal(fl["sample_begin"], of.sample_begin)
    assert np.array_equal(fl["sample_leaves"], of.sample_leaves)


def test_place_mutations_bit_exact(world):
    gm = world.pm.download()
    om = world.om
    say(f"[parity] placed {len(om)} mutations")
    assert len(om) > 50
    assert gm.tobytes() == om.tobytes()


@pytest.mark.parametrize("case", list(cases.SEQ_CASES))
def test_simulate_seq_bit_exact(ctx, world, case):
    kw = cases.SEQ_CASES[case]
    say(f"[parity] simulate_seq case {case}")
    a, d, p, n = core.simulate_seq_arrays(ctx, world.g, world.pf, world.pm, 99, **kw)
    oa, od, op, on = O.simulate_seq(world.lens, world.of, world.om, 99, **kw)
    assert n == on
    assert np.array_equal(d, od)
    assert np.array_equal(a, oa)
    assert np.array_equal(p, op)
    assert d.sum() > 0


def test_replayed_table(ctx, world):
    """A table exported from the oracle and uploaded (replay mode)."""
    say("[parity] replay of an uploaded table")
    pm = core.Mutations.upload(ctx, world.pf, world.om)
    assert pm.download().tobytes() == world.om.tobytes()
    kw = cases.SEQ_CASES["pe_errors"]
    a, d, p, n = core.simulate_seq_arrays(ctx, world.g, world.pf, pm, 5, **kw)
    oa, od, op, on = O.simulate_seq(world.lens, world.of, world.om, 5, **kw)
    assert np.array_equal(a, oa) and np.array_equal(d, od) and np.array_equal(p, op)


def test_dense_windows_use_global_path(ctx):
    """Mutation windows larger than the LDS capacity (1024 records)."""
    say("[parity] dense table: genome 20000, high rates")
    w = build_world(ctx, [20_000], 5, dict(n_cells=400, n_samples=1, sample_size=150, seed=9),
                    dict(snv=2e-4, indel=0.0, pre_indel=0, germline=False, drivers=False, infinite_sites=False), 77)
    assert len(w.om) > 3000
    assert w.pm.download().tobytes() == w.om.tobytes()
    kw = dict(coverage=20, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-2)
    a, d, p, n = core.simulate_seq_arrays(ctx, w.g, w.pf, w.pm, 3, **kw)
    oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, 3, **kw)
    assert np.array_equal(d, od) and np.array_equal(a, oa) and np.array_equal(p, op)
    say("[parity] dense table: materialised reads")
    reads = _reads(ctx, w, 3, kw)
    assert np.array_equal(reads, _oracle_reads(w, 3, kw, len(reads)))


def test_finite_sites_collisions(ctx):
    """infinite_sites_model = FALSE: shared positions and repeated SIDs."""
    say("[parity] finite sites: genome 4000, high rates")
    w = build_world(ctx, [4_000], 6, dict(n_cells=300, n_samples=2, sample_size=60, seed=4),
                    dict(snv=2e-4, indel=2e-5, infinite_sites=False), 31)
    om = w.om
    assert w.pm.download().tobytes() == om.tobytes()
    assert (om["sid_head"] == 0).sum() > 0 or len(np.unique(om["pos"])) < len(om)
    for case in ("pe_errors", "se_errorless", "pe_high_error"):
        kw = cases.SEQ_CASES[case]
        a, d, p, n = core.simulate_seq_arrays(ctx, w.g, w.pf, w.pm, 8, **kw)
        oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, 8, **kw)
        assert np.array_equal(d, od) and np.array_equal(a, oa) and np.array_equal(p, op), case
        reads = _reads(ctx, w, 8, kw)
        assert np.array_equal(reads, _oracle_reads(w, 8, kw, len(reads))), case


def test_batch_equals_plain(ctx, world):
    say("[parity] batched launch of 3 jobs vs oracle")
    spec = [(world, 11, cases.SEQ_CASES["pe_errors"]), (world, 12, cases.SEQ_CASES["purity"]),
            (world, 13, cases.SEQ_CASES["se_errorless"])]
    for cfg in (None, A.LaunchConfig(512, 512, 4, 2, 1), A.LaunchConfig(256, 2048, 2, 16, 0)):
        jobs = [core.SeqJob(ctx, w.g, w.pf, w.pm, seed, **kw) for w, seed, kw in spec]
        for j in jobs:
            for buf in (j.alt, j.depth, j.present):
                buf.fill(0xA5)
        n = core.simulate_seq_batch(ctx, jobs, cfg)
        total = 0
        for (w, seed, kw), j in zip(spec, jobs):
            a, d, p = j.results()
            oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, seed, **kw)
            total += on
            assert np.array_equal(a, oa) and np.array_equal(d, od) and np.array_equal(p, op)
        assert n == total


def test_batch_with_empty_segments(ctx, world):
    """Work items are written on the device from the segment table; segments
    without templates (zero coverage) own no items and share their successor's
    first item, which the item lookup must pass over."""
    say("[parity] batched launch with zero-coverage jobs between others")
    zero = dict(cases.SEQ_CASES["pe_errors"], coverage=0)
    spec = [(world, 21, zero), (world, 22, cases.SEQ_CASES["subset"]), (world, 23, zero),
            (world, 24, cases.SEQ_CASES["pe_high_error"]), (world, 25, zero)]
    for cfg in (None, A.LaunchConfig(256, 512, 3, 8, 0)):
        jobs = [core.SeqJob(ctx, w.g, w.pf, w.pm, seed, **kw) for w, seed, kw in spec]
        for j in jobs:
            for buf in (j.alt, j.depth, j.present):
                buf.fill(0xA5)
        n = core.simulate_seq_batch(ctx, jobs, cfg)
        total = 0
        for (w, seed, kw), j in zip(spec, jobs):
            a, d, p = j.results()
            oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, seed, **kw)
            total += on
            assert np.array_equal(a, oa) and np.array_equal(d, od) and np.array_equal(p, op), (seed, cfg)
        assert n == total


def test_plan_graph_replays(ctx, world):
    """A prepared plan replays its captured hipGraph; every replay, with eager
    calls on the same stream in between and the outputs poisoned before it,
    gives the oracle's tables."""
    say("[parity] batched plan: 3 graph replays with eager work between")
    w = world
    spec = [(11, cases.SEQ_CASES["pe_errors"]), (14, cases.SEQ_CASES["subset"])]
    jobs = [core.SeqJob(ctx, w.g, w.pf, w.pm, seed, **kw) for seed, kw in spec]
    ref = [O.simulate_seq(w.lens, w.of, w.om, seed, **kw) for seed, kw in spec]
    for cfg in (None, A.LaunchConfig(256, 1024, 2, 4, 0)):
        plan = core.SeqPlan(ctx, jobs, cfg)
        for rep in range(3):
            for j in jobs:
                for buf in (j.alt, j.depth, j.present):
                    buf.fill(0x5A)
            n = plan.run()
            assert n == sum(r[3] for r in ref)
            for j, (oa, od, op, on) in zip(jobs, ref):
                a, d, p = j.results()
                assert np.array_equal(a, oa) and np.array_equal(d, od) and np.array_equal(p, op), (cfg, rep)
            # eager work on the same context between replays
            core.simulate_seq_arrays(ctx, w.g, w.pf, w.pm, 99, **cases.SEQ_CASES["short_reads"])
        plan.close()


def _reads(ctx, w, seed, kw):
    out, n = core.simulate_seq_reads(ctx, w.g, w.pf, w.pm, seed, **kw)
    assert n == core.num_reads(w.g, w.pf, **kw)
    return out.numpy()[:n]


def _recount(w, reads, kw, n_out):
    """Count alt/depth from materialised reads (SNV and indel heads)."""
    L = kw["read_size"]
    pe = kw.get("insert_size_mean", 0) > 0
    om = w.om
    M = len(om)
    alt = np.zeros((n_out, M), np.int64)
    dep = np.zeros((n_out, M), np.int64)
    chrs = kw.get("chromosomes") or list(range(len(w.lens)))
    idx = 0
    pos_by_chr = {c: np.where(om["chr"] == c)[0] for c in range(len(w.lens))}
    for so in range(n_out):
        for c in sorted(set(chrs)):
            ln = int(w.lens[c])
            if ln < L:
                continue
            T = int(np.round(kw["coverage"] * ln / (2 * L if pe else L)))
            # llround rounds half away from zero; np.round rounds half to even
            t_exact = kw["coverage"] * ln / (2 * L if pe else L)
            if abs(t_exact - np.floor(t_exact) - 0.5) < 1e-12:
                T = int(np.floor(t_exact)) + 1
            n = T * (2 if pe else 1)
            rows = reads[idx: idx + n]
            idx += n
            ids = pos_by_chr[c]
            mp = om["pos"][ids].astype(np.int64)
            for r in rows:
                x = int(r[0])
                leaf = int(r[1] & 0x3FFFFFF)
                allele = int((r[1] >> 26) & 0xF)
                lo = np.searchsorted(mp, x)
                hi = np.searchsorted(mp, x + L)
                for k in range(lo, hi):
                    q = ids[k]
                    off = int(om["pos"][q]) - x
                    base = int((r[2 + off // 16] >> (2 * (off % 16))) & 3)
                    dep[so, q] += 1
                    if not om["sid_head"][q]:
                        continue
                    if om["kind"][q] == A.PSX_SNV:
                        alt[so, q] += base == om["alt"][q]
                    else:
                        j, car = q, False
                        while j < M and (j == q or not om["sid_head"][j]):
                            rec = om[j]
                            if rec["allele"] == allele:
                                if leaf == w.of.n_leaves:
                                    car |= bool(rec["nature"] == A.PSX_GERMINAL or
                                                (rec["nature"] == A.PSX_PRENEOPLASTIC and
                                                 kw.get("preneoplastic_in_normal", False)))
                                else:
                                    car |= bool(rec["leaf_lo"] <= leaf < rec["leaf_hi"])
                            j += 1
                        alt[so, q] += car
    assert idx == len(reads)
    return alt, dep


def _oracle_reads(w, seed, kw, n):
    return O.simulate_reads(w.lens, w.packed, w.of, w.om, seed, n, **kw)


READ_CASES = dict(cases.SEQ_CASES,
                  # rows too wide for the LDS stage: each lane writes its read in place
                  long_reads=dict(coverage=3, read_size=1000, insert_size_mean=1500, insert_size_stddev=20,
                                  error_rate=4e-3))


@pytest.mark.parametrize("case", list(READ_CASES))
def test_materialised_reads_bit_exact(ctx, world, case):
    """Every read, every bit: start, leaf, allele, mate, error flag, bases."""
    kw = READ_CASES[case]
    say(f"[parity] materialised reads bit-exact, case {case}")
    reads = _reads(ctx, world, 23, kw)
    ref = _oracle_reads(world, 23, kw, len(reads))
    assert reads.shape == ref.shape
    bad = np.nonzero((reads != ref).any(axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:5], reads[bad[:2]], ref[bad[:2]])
    if kw.get("error_rate", 0) > 0:
        assert (reads[:, 1] >> 31).sum() > 0


@pytest.mark.parametrize("case", ["pe_errors", "se_errorless", "purity"])
def test_materialised_reads_match_counts(ctx, world, case):
    kw = cases.SEQ_CASES[case]
    say(f"[parity] materialised reads, case {case}")
    reads = _reads(ctx, world, 21, kw)
    n_out = core.num_out_samples(world.pf, **kw)
    alt, dep = _recount(world, reads, kw, n_out)
    oa, od, op, on = O.simulate_seq(world.lens, world.of, world.om, 21, **kw)
    assert np.array_equal(dep, od)
    assert np.array_equal(alt, oa)


def test_batched_reads_equal_each_job_alone(ctx, world):
    """psx_simulate_seq_reads_batch: job k's rows are the rows the single-job
    call writes for it, for jobs of two forests, SE and PE, purity and
    subsets, at 150 bp (16-byte rows) and 100 bp (any read size)."""
    say("[parity] batched read materialisation: building a second world (30000 bases, 300 cells)")
    w2 = build_world(ctx, [30_000], 5, dict(n_cells=300, n_samples=1, sample_size=50, seed=8), {}, 999)
    subset = dict(coverage=40, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
                  chromosomes=[1], samples=[0], with_normal_sample=False)
    groups = {150: [(world, 23, cases.SEQ_CASES["pe_errors"]), (w2, 9, dict(coverage=12, read_size=150,
                                                                           error_rate=2e-3)),
                    (world, 5, cases.SEQ_CASES["purity"]), (world, 3, subset)],
              100: [(world, 4, cases.SEQ_CASES["se_errorless"]), (w2, 6, dict(coverage=20, read_size=100,
                                                                              insert_size_mean=250,
                                                                              insert_size_stddev=15,
                                                                              error_rate=5e-3))]}
    for L, spec in groups.items():
        say(f"[parity] batched reads, {len(spec)} jobs of {L} bp")
        jobs = [core.SeqJob(ctx, w.g, w.pf, w.pm, seed, **kw) for w, seed, kw in spec]
        out, first, n = core.simulate_seq_reads_batch(ctx, jobs)
        assert n == first[-1] == sum(core.num_reads(w.g, w.pf, **kw) for w, _, kw in spec)
        got = out.numpy()[:n]
        for k, (w, seed, kw) in enumerate(spec):
            alone = _reads(ctx, w, seed, kw)
            assert np.array_equal(got[first[k]:first[k + 1]], alone), (L, k)
        w0, s0, kw0 = spec[0]
        assert np.array_equal(got[first[0]:first[1]], _oracle_reads(w0, s0, kw0, first[1]))
    mixed = [core.SeqJob(ctx, world.g, world.pf, world.pm, 1, **cases.SEQ_CASES["pe_errors"]),
             core.SeqJob(ctx, world.g, world.pf, world.pm, 1, **cases.SEQ_CASES["se_errorless"])]
    with pytest.raises(PsxError, match="one read size"):
        core.simulate_seq_reads_batch(ctx, mixed)
    small = core.DeviceArray(ctx, (10, core.read_words(**cases.SEQ_CASES["pe_errors"])), np.uint32)
    with pytest.raises(PsxError, match="the read buffer holds 10 reads"):
        core.simulate_seq_reads_batch(ctx, mixed[:1], out=small)


# ------------------------------------------------------------ copy numbers

CNA_LENS = np.array([70_000, 33_000], np.uint32)


@pytest.fixture(scope="module")
def cna_world(ctx):
    say("[parity] building copy-number world: genome 70000+33000, 500-cell forest, 5 CNA drivers")
    w = World()
    w.lens = CNA_LENS
    w.packed = O.genome_synth(w.lens, 31)
    w.g = core.Genome(ctx, w.lens, seed=31)
    fd = core.synth_forest(growth=[0.2, 0.45], death=[0.01, 0.01], intro_at=[0, 20], n_cells=500, n_samples=2,
                           sample_size=80, seed=3)
    w.fd = fd
    w.of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    w.pf = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    w.eng = cases.engine(w.packed, w.lens, snv=6e-6, indel=6e-7, pre_snv=20, pre_indel=3)
    w.eng["cna"] = [(0, 0, 5_000, 20_000, 0, A.PSX_CNA_AMP, 2), (0, 1, 1_000, 9_000, 1, A.PSX_CNA_DEL, 0),
                    (1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0), (1, 0, 40_000, 15_000, 1, A.PSX_CNA_AMP, 1),
                    (1, 0, 12_000, 4_000, 0, A.PSX_CNA_DEL, 0)]
    w.om = O.place_mutations(w.lens, w.packed, w.of, 4242, **w.eng)
    w.oev = O.place_cna(w.lens, w.of, **w.eng)
    w.pm = core.Mutations.place(ctx, w.g, w.pf, 4242, **w.eng)
    return w


def test_cna_placement_matches(cna_world):
    w = cna_world
    assert np.array_equal(w.pm.download(), w.om)
    assert w.pm.cna().tobytes() == w.oev.tobytes()
    assert len(w.oev) >= 5


CNA_SEQ = {
    "pe_purity": dict(coverage=30, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
                      purity=0.6),
    "se_pure": dict(coverage=25, read_size=100, error_rate=5e-3, with_normal_sample=False),
    "subset_normal": dict(coverage=20, read_size=120, samples=[1], chromosomes=[0], purity=0.8,
                          preneoplastic_in_normal=True),
    # short reads cover fewer than 48 positions per 64 templates here: errors checked per covered position
    "se_sparse": dict(coverage=40, read_size=30, error_rate=2e-2),
    "pe_sparse": dict(coverage=30, read_size=36, insert_size_mean=100, insert_size_stddev=8, error_rate=1e-2,
                      purity=0.7),
}


@pytest.mark.parametrize("case", sorted(CNA_SEQ))
def test_cna_seq_bit_exact(ctx, cna_world, case):
    say(f"[parity] copy-number sequencing {case}")
    w = cna_world
    kw = CNA_SEQ[case]
    a, d, p, n = core.simulate_seq_arrays(ctx, w.g, w.pf, w.pm, 17, **kw)
    oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, 17, cna=w.oev, **kw)
    assert n == on
    assert np.array_equal(d, od) and np.array_equal(a, oa) and np.array_equal(p, op)


def test_cna_replay_and_batch(ctx, cna_world):
    """A replayed table with its events, batched with a plain job, matches the
    oracle job by job; its reads are the oracle's reads."""
    say("[parity] copy-number replay + batch")
    w = cna_world
    rm = core.Mutations.upload(ctx, w.pf, w.om)
    rm.set_cna(w.oev)
    assert np.array_equal(rm.cna(), w.oev)
    kw = CNA_SEQ["pe_purity"]
    jobs = [core.SeqJob(ctx, w.g, w.pf, rm, 5, **kw),
            core.SeqJob(ctx, w.g, w.pf, core.Mutations.upload(ctx, w.pf, w.om), 6, **kw)]
    n = core.simulate_seq_batch(ctx, jobs)
    oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, 5, cna=w.oev, **kw)
    pa, pd, pp, pn = O.simulate_seq(w.lens, w.of, w.om, 6, **kw)
    a, d, p = jobs[0].results()
    assert np.array_equal(a, oa) and np.array_equal(d, od) and np.array_equal(p, op)
    a, d, p = jobs[1].results()
    assert np.array_equal(a, pa) and np.array_equal(d, pd)
    assert n == on + pn
    rows, nr = core.simulate_seq_reads(ctx, w.g, w.pf, rm, 1, **kw)
    assert nr == core.num_reads_table(ctx, w.g, w.pf, rm, 1, **kw)
    assert np.array_equal(rows.numpy()[:nr], O.simulate_reads(w.lens, w.packed, w.of, w.om, 1, nr, cna=w.oev, **kw))


@pytest.mark.parametrize("case", sorted(CNA_SEQ))
def test_cna_reads_bit_exact(ctx, cna_world, case):
    """Read materialisation with copy-number events: only kept templates have
    reads, in template order; every read equals the oracle's, and their number
    is the count-only call's."""
    say(f"[parity] copy-number read materialisation {case}")
    w = cna_world
    kw = CNA_SEQ[case]
    n = core.num_reads_table(ctx, w.g, w.pf, w.pm, 23, **kw)
    rows, got = core.simulate_seq_reads(ctx, w.g, w.pf, w.pm, 23, **kw)
    assert got == n
    _, _, _, on = O.simulate_seq(w.lens, w.of, w.om, 23, cna=w.oev, **kw)
    assert n == on
    orows = O.simulate_reads(w.lens, w.packed, w.of, w.om, 23, n, cna=w.oev, **kw)
    assert np.array_equal(rows.numpy()[:n], orows)


def test_cna_reads_batched_with_plain_jobs(ctx, world, cna_world):
    """A batched read launch of copy-number and plain jobs: each job's rows are
    its rows alone (and the oracle's)."""
    say("[parity] copy-number read materialisation, batched with plain jobs")
    kw = dict(coverage=20, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=2e-3, purity=0.7)
    spec = [(cna_world, 3), (world, 4), (cna_world, 5)]
    jobs = [core.SeqJob(ctx, w.g, w.pf, w.pm, seed, **kw) for w, seed in spec]
    out, first, n = core.simulate_seq_reads_batch(ctx, jobs)
    got = out.numpy()[:n]
    for k, (w, seed) in enumerate(spec):
        alone, na = core.simulate_seq_reads(ctx, w.g, w.pf, w.pm, seed, **kw)
        assert first[k + 1] - first[k] == na
        assert np.array_equal(got[first[k]:first[k + 1]], alone.numpy()[:na]), k
    w = cna_world
    assert np.array_equal(got[first[0]:first[1]],
                          O.simulate_reads(w.lens, w.packed, w.of, w.om, 3, first[1], cna=w.oev, **kw))


def test_batch_mixes_sparse_and_dense_jobs(ctx, world, cna_world):
    """One batched call with jobs whose reads cover few positions (errors
    checked per covered position) and jobs that cover many (errors per
    template), with and without copy-number events: the call runs one launch
    per kind, and every job matches the oracle."""
    say("[parity] batched launch mixing sparse and dense jobs, plain and copy-number")
    dense = build_world(ctx, [6_000], 8, dict(n_cells=200, n_samples=2, sample_size=50, seed=5),
                        dict(snv=1e-4, indel=1e-5), 41)
    cw = cna_world
    spec = [(world, None, 31, cases.SEQ_CASES["pe_errors"]),               # sparse, plain
            (dense, None, 32, cases.SEQ_CASES["pe_errors"]),               # dense, plain
            (cw, cw.oev, 33, CNA_SEQ["pe_purity"]),                        # dense, copy numbers
            (cw, cw.oev, 34, CNA_SEQ["se_sparse"]),                        # sparse, copy numbers
            (dense, None, 35, cases.SEQ_CASES["pe_high_error"]),           # dense, plain
            (world, None, 36, cases.SEQ_CASES["se_errorless"])]            # no errors
    for cfg in (None, A.LaunchConfig(256, 1024, 2, 8, 0)):
        jobs = [core.SeqJob(ctx, w.g, w.pf, w.pm, seed, **kw) for w, ev, seed, kw in spec]
        for j in jobs:
            for buf in (j.alt, j.depth, j.present):
                buf.fill(0xA5)
        n = core.simulate_seq_batch(ctx, jobs, cfg)
        total = 0
        for (w, ev, seed, kw), j in zip(spec, jobs):
            a, d, p = j.results()
            oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, seed, cna=ev, **kw)
            total += on
            assert np.array_equal(d, od) and np.array_equal(a, oa) and np.array_equal(p, op), (seed, cfg)
        assert n == total


def test_empty_table_and_zero_coverage(ctx):
    """A mutation table with no records (every read is the reference) and a
    run with zero coverage (no reads), in both modes."""
    say("[parity] empty mutation table; zero coverage")
    w = build_world(ctx, [9_000, 3_000], 12, dict(n_cells=60, n_samples=1, sample_size=20, seed=2),
                    dict(snv=0.0, indel=0.0, pre_snv=0, pre_indel=0, drivers=False, germline=False), 5)
    assert len(w.om) == 0 and w.pm.n == 0
    kw = cases.SEQ_CASES["pe_errors"]
    a, d, p, n = core.simulate_seq_arrays(ctx, w.g, w.pf, w.pm, 7, **kw)
    oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, 7, **kw)
    assert n == on and n > 0 and a.size == 0 and d.size == 0
    reads = _reads(ctx, w, 7, kw)
    assert np.array_equal(reads, _oracle_reads(w, 7, kw, len(reads)))
    zero = dict(kw, coverage=0)
    a, d, p, n = core.simulate_seq_arrays(ctx, w.g, w.pf, w.pm, 7, **zero)
    assert n == 0
    out, got = core.simulate_seq_reads(ctx, w.g, w.pf, w.pm, 7, **zero)
    assert got == 0


def test_longest_reads_and_ragged_chromosomes(ctx):
    """Reads of the largest supported size (4096 bases): a chromosome exactly
    one read long (one start), one shorter than a read (no reads), one of a
    length that is no multiple of 16; single and paired ends at the largest
    reach; both modes."""
    say("[parity] 4096-base reads on chromosomes of 4096, 4095 and 5003 bases")
    w = build_world(ctx, [4_096, 4_095, 5_003], 14, dict(n_cells=80, n_samples=1, sample_size=30, seed=6),
                    dict(snv=3e-4, indel=3e-5), 9)
    for kw in (dict(coverage=30, read_size=4096, error_rate=1e-3),
               dict(coverage=12, read_size=2000, insert_size_mean=3000, insert_size_stddev=30, error_rate=2e-3)):
        a, d, p, n = core.simulate_seq_arrays(ctx, w.g, w.pf, w.pm, 10, **kw)
        oa, od, op, on = O.simulate_seq(w.lens, w.of, w.om, 10, **kw)
        assert n == on and n > 0
        assert np.array_equal(d, od) and np.array_equal(a, oa) and np.array_equal(p, op), kw
        reads = _reads(ctx, w, 10, kw)
        assert np.array_equal(reads, _oracle_reads(w, 10, kw, len(reads))), kw


def _expected_vaf32(om, a, d, p):
    """The table's VAF rule per record in fp32 (numpy's IEEE division)."""
    S, M = a.shape
    head = om["sid_head"] != 0
    exp = np.zeros((S, M), np.float32)
    heads = np.nonzero(head)[0]
    ends = np.append(heads[1:], M)
    with np.errstate(divide="ignore", invalid="ignore"):
        for i, e in zip(heads, ends):
            held = (a[:, i] > 0) | (p[:, i:e] != 0).any(axis=1)
            exp[held, i] = np.float32(a[held, i]) / np.float32(d[held, i])
    return exp


def test_vaf_f32_on_device(ctx, world):
    """The result table's VAF in fp32 on the device: bit for bit the fp32
    quotient (NaN where a held SID has no depth), 0 off SID heads and where
    the sample does not hold the SID; within one fp32 rounding of the
    double-precision table of psx_seq_table. Infinite and finite sites (SIDs
    of several records)."""
    say("[parity] fp32 VAF on the device")
    fin = build_world(ctx, [4_000], 6, dict(n_cells=300, n_samples=2, sample_size=60, seed=4),
                      dict(snv=2e-4, indel=2e-5, infinite_sites=False), 31)
    for w, kw in ((world, cases.SEQ_CASES["purity"]), (fin, cases.SEQ_CASES["pe_high_error"])):
        S = core.num_out_samples(w.pf, **kw)
        M = w.pm.n
        alt = core.DeviceArray(ctx, (S, M), np.int32)
        dep = core.DeviceArray(ctx, (S, M), np.int32)
        pres = core.DeviceArray(ctx, (S, M), np.uint8)
        core.simulate_seq_device(ctx, w.g, w.pf, w.pm, 17, alt, dep, pres, **kw)
        vd = core.seq_vaf_f32(ctx, w.pm, alt, dep, pres)
        vaf = vd.numpy()[:, :M]
        a, d, p = alt.numpy(), dep.numpy(), pres.numpy()
        for x in (alt, dep, pres, vd):
            x.close()
        exp = _expected_vaf32(w.om, a, d, p)
        assert np.array_equal(np.isnan(vaf), np.isnan(exp))
        ok = ~np.isnan(exp)
        assert np.array_equal(vaf[ok].view(np.uint32), exp[ok].view(np.uint32))
        assert (vaf[ok] > 0).sum() > 0
        rows, occ, cov, v64 = core.seq_table(w.om, a, d, p, include_non_sequenced=True)
        v32 = vaf[:, rows.astype(np.int64)]
        fin64 = ~np.isnan(v64)
        assert np.array_equal(np.isnan(v32), ~fin64)
        assert np.all(np.abs(v32[fin64].astype(np.float64) - v64[fin64]) <= 2.0 ** -24 * np.abs(v64[fin64]))
