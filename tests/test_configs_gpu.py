"""The HIP path on the benchmark's batched, copy-number and serving workloads
at their own sizes, and on the API paths and ploidies the other GPU tests do
not reach, each against the oracle and, where the work is shared between
jobs, against every job run alone:

* the copy-number + purity workload (1e5-cell forest, one 1e4-cell sample and
  the normal sample, 100 Mb over 10 chromosomes, allele-specific segments and
  a whole-genome doubling, purity 0.6, 500x): the oracle on the two shortest
  chromosomes gives exactly their rows of the full GPU table;
* the batched workload (256 jobs over a shared 50 Mb reference): one
  hipGraph-captured plan and the autotuner's configuration give every job's
  tables exactly as the job run alone, and the oracle's for the smallest jobs;
* the serving pool (8 resident forests of 1e3-1e6 cells, 50 Mb): a burst of
  requests through the continuous-batching server, each equal to the request
  alone, the smallest equal to the oracle;
* ploidy 3 and 4, with and without copy-number events; the normal sample
  alone (simulate_normal_seq); a FACS-labelled run (cell_labelling).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import cases
import process_amd as pa
from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oforest(fd):
    return O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])


def same(got, exp, what=""):
    a, d, p, n = got
    oa, od, op, on = exp
    S, M = a.shape
    assert n == on, (what, n, on)
    assert np.array_equal(d, np.reshape(od, (S, M))), what
    assert np.array_equal(a, np.reshape(oa, (S, M))), what
    assert np.array_equal(p, np.reshape(op, (S, M))), what


# ---------------------------------------------------------------- ploidy

@pytest.mark.parametrize("ploidy", [3, 4])
@pytest.mark.parametrize("with_cna", [False, True])
def test_ploidy_against_oracle(ctx, ploidy, with_cna):
    lens = np.array([70_000, 41_000], np.uint32)
    packed = O.genome_synth(lens, 13)
    g = core.Genome(ctx, lens, seed=13)
    fd = cases.forest(n_cells=700, n_samples=2, sample_size=120, seed=8)
    f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    of = oforest(fd)
    eng = cases.engine(packed, lens, snv=4e-6, indel=4e-7, pre_snv=6, pre_indel=2)
    eng["ploidy"] = ploidy
    eng["drivers"] = [(s, c, p, a, -1) for s, c, p, a, _ in eng["drivers"]]
    eng["germline"] = [(c, p, a, min(al + 1, ploidy - 1)) for c, p, a, al in eng["germline"]]
    if with_cna:
        eng["cna"] = [(0, 0, 5_000, 30_000, ploidy - 1, A.PSX_CNA_AMP, 2), (0, 1, 1_000, 8_000, 1, A.PSX_CNA_DEL, 0),
                      (1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0), (1, 1, 20_000, 15_000, 0, A.PSX_CNA_AMP, 1)]
    say(f"[configs] ploidy {ploidy}, copy numbers {with_cna}")
    m = core.Mutations.place(ctx, g, f, 21, **eng)
    om = O.place_mutations(lens, packed, of, 21, **eng)
    assert m.download().tobytes() == om.tobytes()
    assert np.all(om["allele"] < ploidy) and len(set(om["allele"].tolist())) == ploidy
    ev = m.cna()
    if with_cna:
        assert ev.tobytes() == O.place_cna(lens, of, **eng).tobytes() and len(ev) > 0
    for seq in (dict(coverage=35, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=2e-3,
                     purity=0.7, preneoplastic_in_normal=True),
                dict(coverage=20, read_size=100, error_rate=1e-3)):
        got = core.simulate_seq_arrays(ctx, g, f, m, 33, **seq)
        exp = O.simulate_seq(lens, of, om, 33, ploidy=ploidy, cna=ev if with_cna else None, **seq)
        same(got, exp, (ploidy, with_cna, seq))
        assert got[1].sum() > 0


# ------------------------------------------------------- normal-only, FACS

def api_build(seed=3):
    sig = cases.signatures(2, 11)
    ref = pa.ReferenceGenome({"21": 60_000, "22": 45_000}, seed=17)
    eng = pa.MutationEngine(ref, {"SBS1": sig[0], "SBS5": sig[1]}, germline=[("22", 1001, "T", 1), ("21", 501, "G", 0)],
                            context_sampling=1)
    eng.add_mutant("A", {"SNV": 5e-6, "indel": 5e-7})
    eng.add_mutant("B", {"SNV": 8e-6, "indel": 5e-7})
    eng.add_exposure({"SBS1": 0.7, "SBS5": 0.3})
    sf = pa.SamplesForest.synthetic(mutants={"A": (0.2, 0.01), "B": (0.3, 0.01)}, introductions={"B": 100},
                                    n_cells=600, samples={"S1": 120, "S2": 120}, seed=seed)
    return eng, sf


def table_against_oracle(df, names, om, oa, od, op, include=False):
    rows, occ, cov, vaf = O.seq_table(om, oa, od, op, include)
    assert len(df) == len(rows)
    assert (df["chr_pos"].to_numpy() == om["pos"][rows.astype(np.int64)] + 1).all()
    for s, name in enumerate(names):
        assert np.array_equal(df[f"{name}.occurrences"].to_numpy(), occ[s]), name
        assert np.array_equal(df[f"{name}.coverage"].to_numpy(), cov[s]), name
        assert np.allclose(df[f"{name}.VAF"].to_numpy(), vaf[s], equal_nan=True), name


@pytest.mark.parametrize("with_preneoplastic", [False, True])
def test_normal_only_against_oracle(ctx, with_preneoplastic):
    """simulate_normal_seq (src/seq_simulation.cpp:603-679): the normal sample
    alone at purity 1, with and without the pre-neoplastic mutations."""
    eng, sf = api_build()
    pf = eng.place_mutations(sf, 8, 2, seed=5)
    res = pa.simulate_normal_seq(pf, sequencer=pa.BasicIlluminaSequencer(2e-3), coverage=60, insert_size_mean=350,
                                 with_preneoplastic=with_preneoplastic, write_SAM=False,
                                 include_non_sequenced_mutations=True, seed=9)
    om = pf.records()
    of = O.Forest(sf.cell_id, sf.ancestor, sf.species_index(), sf.sample_index(sf.sample_names), sf.birth_time, 2)
    oa, od, op, on = O.simulate_seq(eng.reference.lengths, of, om, 9, coverage=60, read_size=150,
                                    insert_size_mean=350, insert_size_stddev=10, error_rate=2e-3, samples=[],
                                    with_normal_sample=True, purity=1.0,
                                    preneoplastic_in_normal=with_preneoplastic)
    df = res["mutations"]
    table_against_oracle(df, ["normal_sample"], om, oa, od, op, include=True)
    # the normal genome holds the germline (and the pre-neoplastic mutations when asked, at about half the
    # reads); rows of other SIDs come from sequencing errors that read their alternative base, a read or two
    vaf = df["normal_sample.VAF"].to_numpy()
    germ = (df["classes"] == "germinal").to_numpy()
    pre = (df["classes"] == "preneoplastic").to_numpy()
    passenger = (df["classes"] == "passenger").to_numpy()
    assert germ.any() and (vaf[germ] > 0.2).all()
    assert (vaf[passenger] < 0.1).all()
    if with_preneoplastic:
        assert pre.any() and (vaf[pre] > 0.2).any()
    else:
        assert (vaf[pre] < 0.1).all()


def test_facs_labelled_against_oracle(ctx):
    """cell_labelling (src/seq_simulation.cpp:183-243): every sample splits by
    the label of its cells into "<sample>_<label>" output samples, in order
    of first appearance; the oracle run on the relabelled forest gives the
    same table."""
    eng, sf = api_build(seed=7)
    pf = eng.place_mutations(sf, 4, 1, seed=2)

    def label(c):
        return "" if c.mutant == "A" and c.birth_time < 25 else ("B" if c.mutant == "B" else "late")

    res = pa.simulate_seq(pf, sequencer=pa.BasicIlluminaSequencer(1e-3), coverage=30, insert_size_mean=350,
                          cell_labelling=label, purity=0.9, include_non_sequenced_mutations=True, seed=12)
    df = res["mutations"]
    names, index = [], {}
    new_sample = np.full(len(sf.cell_id), -1, np.int32)
    for s in sf.sample_names:
        for i in np.nonzero(sf.sample == s)[0]:
            lab = label(pa.api.SampledCell(int(sf.cell_id[i]), str(sf.mutant[i]), str(sf.epistate[i]),
                                           str(sf.mutant[i]), float(sf.birth_time[i])))
            n = s if lab == "" else f"{s}_{lab}"
            if n not in index:
                index[n] = len(names)
                names.append(n)
            new_sample[i] = index[n]
    assert len(names) >= 3
    got = sorted(c[:-4] for c in df.columns if c.endswith(".VAF"))
    assert got == sorted(names + ["normal_sample"])
    om = pf.records()
    of = O.Forest(sf.cell_id, sf.ancestor, sf.species_index(), new_sample, sf.birth_time, len(names))
    oa, od, op, on = O.simulate_seq(eng.reference.lengths, of, om, 12, coverage=30, read_size=150,
                                    insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3, purity=0.9)
    table_against_oracle(df, names + ["normal_sample"], om, oa, od, op, include=True)


# -------------------------------------------------- copy numbers + purity

def test_cna_purity_workload_against_oracle(ctx):
    """The full copy-number + purity workload on the GPU; its two shortest
    chromosomes' rows against the oracle run on them alone (a chromosome's
    counts depend on that chromosome only, copy-number events included)."""
    w = workloads.cna_purity()
    say("[configs] copy-number workload: 1e5-cell forest, 100 Mb / 10 chromosomes, 500x, purity 0.6")
    b = workloads.build(ctx, w)
    ev = b.muts.cna()
    assert len(ev) > 0 and (ev["kind"] == A.PSX_CNA_WGD).any()
    a, d, p, n = core.simulate_seq_arrays(ctx, b.genome, b.forest, b.muts, w.seq_seed, **w.seq)
    say(f"[configs] {b.muts.n} mutations, {len(ev)} copy-number events, {n} accepted reads")
    assert n > 500_000_000
    lens = np.asarray(w.lengths, np.uint32)
    chosen = [int(c) for c in np.argsort(lens)[:2]]
    recs = b.muts.download()
    of = oforest(w.forest)
    say(f"[configs] oracle on chromosomes {chosen} ({int(lens[chosen].sum())} bases)")
    oa, od, op, on = O.simulate_seq(lens, of, recs, w.seq_seed, cna=ev, chromosomes=chosen, **w.seq)
    rows = np.isin(recs["chr"], chosen)
    S, M = a.shape
    oa, od, op = (x.reshape(S, M) for x in (oa, od, op))
    assert rows.sum() > 100 and on > 1_000_000
    assert np.array_equal(d[:, rows], od[:, rows])
    assert np.array_equal(a[:, rows], oa[:, rows])
    assert np.array_equal(p[:, rows], op[:, rows])
    # the events touch the chosen chromosomes (or the doubling covers every chromosome)
    assert ((np.isin(ev["chr"], chosen)) | (ev["chr"] == A.PSX_ALL_CHROMOSOMES)).any()


# ---------------------------------------------------------------- batched

def test_batched_workload_256_jobs(ctx):
    """All 256 jobs of the batched workload in one hipGraph-captured plan and
    with the autotuner's configuration: each job's tables equal the job run
    alone; the three smallest jobs' equal the oracle's."""
    say("[configs] batched workload: 256 jobs over a shared 50 Mb reference")
    B = workloads.build_batched(ctx, n_jobs=256)
    lens, _, specs = workloads.batched_jobs(seed=99, n_jobs=256)
    plan = core.SeqPlan(ctx, B.jobs)
    n_plan = plan.run()
    say(f"[configs] plan replay: {n_plan} reads in {ctx.last_kernel_ms():.1f} ms")
    assert n_plan == B.reads > 5e9
    plan_res = [j.results() for j in B.jobs]
    plan.close()
    say("[configs] every job alone")
    for k, j in enumerate(B.jobs):
        a, d, p, n = core.simulate_seq_arrays(ctx, j.genome, j.forest, j.muts, j.seed, **j.seq)
        pa_, pd_, pp_ = plan_res[k]
        assert np.array_equal(a, pa_) and np.array_equal(d, pd_) and np.array_equal(p, pp_), k
    cfg, ms = core.autotune(ctx, B.jobs)
    say(f"[configs] tuned: block {cfg.block_threads}, lds {cfg.lds_records}, strata {cfg.strata_per_item}, "
        f"blocks/CU {cfg.blocks_per_cu} ({ms:.1f} ms)")
    for j in B.jobs:
        for buf in (j.alt, j.depth, j.present):
            buf.fill(0x5A)
    cfg.use_graph = 1
    n_tuned = core.simulate_seq_batch(ctx, B.jobs, cfg)
    assert n_tuned == n_plan
    for k, j in enumerate(B.jobs):
        a, d, p = j.results()
        assert np.array_equal(a, plan_res[k][0]) and np.array_equal(d, plan_res[k][1]), k
        assert np.array_equal(p, plan_res[k][2]), k
    reads = [core.num_reads(j.genome, j.forest, **j.seq) for j in B.jobs]
    for k in np.argsort(reads)[:3]:
        s, j = specs[k], B.jobs[k]
        fd = core.synth_forest(growth=[0.2], death=[0.01], intro_at=[0], n_cells=2 * s["leaves"], n_samples=1,
                               sample_size=s["leaves"], seed=s["forest_seed"])
        say(f"[configs] oracle on job {k} ({s['leaves']} leaves, {reads[k]} reads)")
        exp = O.simulate_seq(np.asarray(lens, np.uint32), oforest(fd), j.muts.download(), j.seed, **j.seq)
        same((*plan_res[k], reads[k]), exp, k)


# ---------------------------------------------------------------- serving

def test_serving_pool_burst(ctx):
    """A burst of the serving workload's requests (8 resident forests of
    1e3-1e6 cells over 50 Mb, coverage 30-500x, random sample subsets)
    through the continuous-batching server: every result equals the request
    run alone, and the request with the fewest reads equals the oracle."""
    say("[configs] serving pool: 8 forests of 1e3-1e6 cells over 50 Mb")
    pool = workloads.build_serving_pool(ctx)
    reqs = workloads.serving_requests(pool, 48, rate=0.0, seed=21)
    alone = []
    for _, k, seed, seq in reqs:
        alone.append(core.simulate_seq_arrays(ctx, pool.genome, pool.forests[k], pool.muts[k], seed, **seq))
    srv = core.Server(ctx, 4 << 30, max_inflight=2, max_batch=16)
    try:
        tickets = [srv.submit(pool.genome, pool.forests[k], pool.muts[k], seed, **seq) for _, k, seed, seq in reqs]
        batches = []
        for i, t in enumerate(tickets):
            a, d, p, st = srv.wait(t, timeout_ms=120_000)
            assert np.array_equal(a, alone[i][0]) and np.array_equal(d, alone[i][1]), i
            assert np.array_equal(p, alone[i][2]) and st["n_reads"] == alone[i][3], i
            batches.append(st["batch_size"])
        tot = srv.totals()
    finally:
        srv.close()
    say(f"[configs] served {tot['completed']} requests in {tot['launches']} launches (largest batch {max(batches)})")
    assert tot["completed"] == len(reqs) and max(batches) > 1
    i = int(np.argmin([x[3] for x in alone]))
    _, k, seed, seq = reqs[i]
    lens = np.asarray(pool.genome.lengths, np.uint32)
    of = oforest(pool.descs[k])
    say(f"[configs] oracle on request {i} ({alone[i][3]} reads)")
    exp = O.simulate_seq(lens, of, pool.muts[k].download(), seed, **seq)
    same(alone[i], exp, i)


# ----------------------------------------------------------- passenger CNAs

def test_passenger_cnas_against_oracle(ctx):
    """Passenger CNAs at per-species rates drawn from a catalogue, with
    copy-number drivers (one with a drawn allele) and a doubling: the placed
    events and mutations equal the oracle's, and so does sequencing over
    them."""
    lens = np.array([120_000, 64_000], np.uint32)
    packed = O.genome_synth(lens, 5)
    g = core.Genome(ctx, lens, seed=5)
    fd = core.synth_forest(growth=[0.2, 0.45], death=[0.01, 0.01], intro_at=[0, 25], n_cells=800, n_samples=2,
                           sample_size=120, seed=19)
    f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    of = oforest(fd)
    eng = cases.engine(packed, lens, snv=3e-6, indel=3e-7, pre_snv=5, pre_indel=1)
    eng["cna"] = [(0, 0, 2_000, 30_000, A.PSX_CNA_RANDOM_ALLELE, A.PSX_CNA_AMP, 1),
                  (1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0), (1, 1, 9_000, 20_000, 1, A.PSX_CNA_DEL, 0)]
    eng["cna_rate"] = [4e-8, 1e-7]
    eng["cna_catalogue"] = [(0, 50_000, 40_000, A.PSX_CNA_AMP), (0, 5_000, 8_000, A.PSX_CNA_DEL),
                            (1, 30_000, 25_000, A.PSX_CNA_AMP), (1, 1_000, 3_000, A.PSX_CNA_DEL)]
    m = core.Mutations.place(ctx, g, f, 314, **eng)
    ev = m.cna()
    oev = O.place_cna(lens, of, seed=314, **eng)
    assert ev.tobytes() == oev.tobytes()
    n_drv = len(O.place_cna(lens, of, seed=314, **dict(eng, cna_rate=None)))
    say(f"[configs] passenger CNAs: {len(ev) - n_drv} passenger and {n_drv} driver events")
    assert len(ev) - n_drv > 10
    om = O.place_mutations(lens, packed, of, 314, **eng)
    assert m.download().tobytes() == om.tobytes()
    seq = dict(coverage=40, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3, purity=0.8)
    got = core.simulate_seq_arrays(ctx, g, f, m, 8, **seq)
    same(got, O.simulate_seq(lens, of, om, 8, cna=oev, **seq), "passenger CNAs")


def test_many_events_per_chromosome(ctx):
    """Short passenger CNAs at a high rate: 70 events on chromosome 0 (more
    than the 64 a stratum's event mask covers, so every template scans the
    whole list) and 60 on chromosome 1 (masked per stratum). Events, count
    tables and every materialised read equal the oracle's."""
    lens = np.array([120_000, 64_000], np.uint32)
    packed = O.genome_synth(lens, 5)
    g = core.Genome(ctx, lens, seed=5)
    fd = core.synth_forest(growth=[0.2, 0.45], death=[0.01, 0.01], intro_at=[0, 25], n_cells=800, n_samples=2,
                           sample_size=120, seed=19)
    f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    of = oforest(fd)
    eng = cases.engine(packed, lens, snv=3e-6, indel=3e-7, pre_snv=5, pre_indel=1)
    eng["cna"] = [(0, 0, 2_000, 30_000, A.PSX_CNA_RANDOM_ALLELE, A.PSX_CNA_AMP, 1), (1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0)]
    eng["cna_rate"] = [4e-7, 8e-7]
    eng["cna_catalogue"] = [(0, 50_000, 4_000, A.PSX_CNA_AMP), (0, 5_000, 800, A.PSX_CNA_DEL),
                            (0, 80_000, 30_000, A.PSX_CNA_DEL), (1, 30_000, 2_500, A.PSX_CNA_AMP),
                            (1, 1_000, 3_000, A.PSX_CNA_DEL)]
    m = core.Mutations.place(ctx, g, f, 314, **eng)
    ev = m.cna()
    oev = O.place_cna(lens, of, seed=314, **eng)
    assert ev.tobytes() == oev.tobytes()
    every = (ev["chr"] == A.PSX_ALL_CHROMOSOMES).sum()
    assert (ev["chr"] == 0).sum() + every > 64 and (ev["chr"] == 1).sum() + every <= 64
    om = O.place_mutations(lens, packed, of, 314, **eng)
    assert m.download().tobytes() == om.tobytes()
    seq = dict(coverage=40, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3, purity=0.8)
    say("[configs] 130 copy-number events: counts")
    exp = O.simulate_seq(lens, of, om, 8, cna=oev, **seq)
    same(core.simulate_seq_arrays(ctx, g, f, m, 8, **seq), exp, "many events")
    say("[configs] 130 copy-number events: reads")
    rows, n = core.simulate_seq_reads(ctx, g, f, m, 8, **seq)
    assert n == exp[3]
    assert np.array_equal(rows.numpy()[:n], O.simulate_reads(lens, packed, of, om, 8, n, cna=oev, **seq))


def test_cna_drivers_through_add_mutant(ctx):
    """MutationEngine$add_mutant with an amplification, a deletion and a
    whole-genome doubling among the drivers, and a passenger CNA rate over a
    catalogue file: the placed events equal the oracle's for the same
    parameters, and simulate_seq's table equals the oracle's."""
    import os
    import tempfile
    sig = cases.signatures(1, 3)
    ref = pa.ReferenceGenome({"1": 80_000, "2": 50_000}, seed=23)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cnas.tsv")
        with open(path, "w") as fh:
            fh.write("chr\tbegin\tend\tmajor\tminor\ttype\tstudy\n1\t1001\t9000\t2\t1\tT\tS\n"
                     "2\t20001\t26000\t1\t0\tT\tS\n1\t40001\t41000\t3\t1\tU\tS\n")
        eng = pa.MutationEngine(ref, {"SBS1": sig[0]}, passenger_CNAs=path, tumour_type="T", context_sampling=1)
    assert eng.passenger_CNAs == [("1", 1001, 9000, "A"), ("2", 20001, 26000, "D")]
    eng.add_mutant("A", {"SNV": 5e-6, "indel": 2e-7, "CNA": 5e-8})
    eng.add_mutant("B", {"SNV": 8e-6, "CNA": 1e-7},
                   drivers=[pa.Amplification("1", 30_001, 10_000, src_allele=1), pa.WGD,
                            pa.Deletion("2", 5_001, 7_000, allele=0)])
    eng.add_exposure({"SBS1": 1.0})
    sf = pa.SamplesForest.synthetic(mutants={"A": (0.2, 0.01), "B": (0.4, 0.01)}, introductions={"B": 60},
                                    n_cells=700, samples={"S1": 100, "S2": 100}, seed=12)
    pf = eng.place_mutations(sf, 3, 0, seed=99)
    params = eng._params(sf, 3, "SBS1", 0)
    of = O.Forest(sf.cell_id, sf.ancestor, sf.species_index(), sf.sample_index(sf.sample_names), sf.birth_time, 2)
    oev = O.place_cna(ref.lengths, of, seed=99, **params)
    ev = pf.mutations.cna()
    assert ev.tobytes() == oev.tobytes()
    kinds = set(ev["kind"].tolist())
    assert kinds == {A.PSX_CNA_AMP, A.PSX_CNA_DEL, A.PSX_CNA_WGD}
    om = pf.records()
    assert om.tobytes() == O.place_mutations(ref.lengths, eng._packed, of, 99, **params).tobytes()
    res = pa.simulate_seq(pf, sequencer=pa.BasicIlluminaSequencer(1e-3), coverage=30, insert_size_mean=350,
                          purity=0.7, include_non_sequenced_mutations=True, seed=4)
    oa, od, op, on = O.simulate_seq(ref.lengths, of, om, 4, cna=oev, coverage=30, read_size=150,
                                    insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3, purity=0.7)
    table_against_oracle(res["mutations"], ["S1", "S2", "normal_sample"], om, oa, od, op, include=True)
