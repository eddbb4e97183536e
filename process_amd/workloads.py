"""Seeded synthetic workloads: the plain, large, batched and serving
configurations of the benchmark, built from the library's synthetic inputs
(the reference's FASTA, COSMIC signatures and spatial simulation need
network access or stay on the CPU)."""
from __future__ import annotations

from dataclasses import dataclass, field

import time

import numpy as np

from . import core


def log_uniform_lengths(n: int, total: int, lo: float, hi: float, rng) -> np.ndarray:
    """n chromosome lengths, log-uniform on [lo, hi] before normalising to total."""
    w = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    lens = np.floor(w / w.sum() * total).astype(np.int64)
    lens[np.argmax(lens)] += total - lens.sum()
    return lens.astype(np.uint32)


def dirichlet_signature(rng) -> np.ndarray:
    return rng.dirichlet(np.ones(96))


@dataclass
class Workload:
    name: str
    lengths: np.ndarray
    genome_seed: int
    forest: dict
    engine: dict
    seq: dict
    place_seed: int
    seq_seed: int
    notes: dict = field(default_factory=dict)


def base_at(genome: core.Genome, packed_cache: dict, chr_: int, pos: int) -> int:
    if "packed" not in packed_cache:
        packed_cache["packed"] = genome.download()
    lens = genome.lengths
    off = int(sum((int(l) + 15) // 16 for l in lens[:chr_]))
    return int((packed_cache["packed"][off + pos // 16] >> (2 * (pos % 16))) & 3)


def plain_large(seed: int = 2026, n_cells: int = 1_000_000, sample_size: int = 10_000,
                total_bases: int = 300_000_000) -> Workload:
    """Two species (growth 0.2; driver 0.3 introduced at 1,000 cells) grown to
    1e6 cells, three samples of ~1e4 cells, 300 Mb over 22 chromosomes with
    log-uniform lengths, SNV 1e-8 and indel 1e-9 per base per division,
    200x paired-end 150 bp reads (insert 350 +- 10), error rate 1e-3, normal
    sample included."""
    rng = np.random.default_rng(seed)
    lens = log_uniform_lengths(22, total_bases, 1.0, 10.0, rng)
    forest = core.synth_forest(growth=[0.2, 0.3], death=[0.01, 0.01], intro_at=[0, 1000], n_cells=n_cells,
                               n_samples=3, sample_size=sample_size, seed=seed + 1)
    sig = dirichlet_signature(rng)
    engine = dict(snv_rate=[1e-8, 1e-8], indel_rate=[1e-9, 1e-9], sbs=sig[None, :], epoch_start=[0.0],
                  exposure=[[1.0]], n_pre_snv=0, n_pre_indel=0, pre_sbs=0, indel_max_len=10, indel_len_p=0.5,
                  infinite_sites=True, ploidy=2, drivers=[], germline=[])
    seq = dict(coverage=200.0, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
               purity=1.0, with_normal_sample=True)
    return Workload("plain_large", lens, seed + 2, forest, engine, seq, seed + 3, seed + 4,
                    notes=dict(n_cells=n_cells, samples=3, sample_size=sample_size, genome=int(lens.sum()),
                               chromosomes=22, driver_species=1, driver_intro_at=1000))


def plain_small(seed: int = 7) -> Workload:
    """One species (growth 0.2, death 0.01) grown to ~5,000 cells, one sample
    of 2,500 cells, a 10 Mb single-chromosome reference, SNV 1e-8, 80x
    paired-end 150 bp reads, error rate 1e-3, normal sample included."""
    rng = np.random.default_rng(seed)
    lens = np.array([10_000_000], np.uint32)
    forest = core.synth_forest(growth=[0.2], death=[0.01], intro_at=[0], n_cells=5000, n_samples=1,
                               sample_size=2500, seed=seed + 1)
    sig = dirichlet_signature(rng)
    engine = dict(snv_rate=[1e-8], indel_rate=[0.0], sbs=sig[None, :], epoch_start=[0.0], exposure=[[1.0]],
                  n_pre_snv=0, n_pre_indel=0, pre_sbs=0, indel_max_len=10, indel_len_p=0.5, infinite_sites=True,
                  ploidy=2, drivers=[], germline=[])
    seq = dict(coverage=80.0, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
               purity=1.0, with_normal_sample=True)
    return Workload("plain_small", lens, seed + 2, forest, engine, seq, seed + 3, seed + 4)


def add_driver(w: Workload, genome: core.Genome, species: int = 1) -> None:
    """A driver SNV of `species` at the middle of chromosome 0."""
    cache: dict = {}
    pos = int(w.lengths[0]) // 2
    ref = base_at(genome, cache, 0, pos)
    w.engine["drivers"] = [(species, 0, pos, (ref + 1) % 4, -1)]


@dataclass
class Built:
    workload: Workload
    genome: core.Genome
    forest: core.Forest
    muts: core.Mutations
    place_ms: float


def build(ctx: core.Context, w: Workload, driver_species: int | None = None) -> Built:
    g = core.Genome(ctx, w.lengths, seed=w.genome_seed)
    if driver_species is not None:
        add_driver(w, g, driver_species)
    fd = w.forest
    f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    m = core.Mutations.place(ctx, g, f, w.place_seed, **w.engine)
    return Built(w, g, f, m, ctx.last_place_ms())


def batched_jobs(seed: int = 99, n_jobs: int = 256, total_bases: int = 50_000_000, min_cells: int = 1_000,
                 max_cells: int = 100_000, min_cov: float = 30.0, max_cov: float = 300.0):
    """Independent jobs over one shared reference: leaves log-uniform in
    [min_cells, max_cells], coverage log-uniform in [min_cov, max_cov]."""
    rng = np.random.default_rng(seed)
    lens = log_uniform_lengths(4, total_bases, 1.0, 4.0, rng)
    jobs = []
    for k in range(n_jobs):
        leaves = int(np.exp(rng.uniform(np.log(min_cells), np.log(max_cells))))
        cov = float(np.exp(rng.uniform(np.log(min_cov), np.log(max_cov))))
        jobs.append(dict(leaves=leaves, coverage=cov, forest_seed=int(rng.integers(1 << 31)),
                         place_seed=int(rng.integers(1 << 31)), seq_seed=int(rng.integers(1 << 31))))
    sig = dirichlet_signature(rng)
    return lens, sig, jobs


@dataclass
class BatchedSet:
    genome: core.Genome
    jobs: list          # core.SeqJob
    specs: list         # dicts of batched_jobs
    forests: list
    muts: list
    place_ms: float       # HIP-event span of the batched placement
    reads: int
    place_wall_ms: float = 0.0  # wall clock of the batched placement call
    engine: dict = field(default_factory=dict)  # the engine parameters every job was placed with


def seq_job_params(coverage: float) -> dict:
    return dict(coverage=coverage, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
                purity=1.0, with_normal_sample=True)


def build_batched(ctx: core.Context, seed: int = 99, n_jobs: int = 256, **kw) -> BatchedSet:
    """The batched workload: n_jobs one-species forests (growth 0.2, death
    0.01) grown to twice their sample size, one sample each, over a shared
    50 Mb reference of 4 chromosomes; SNV 1e-8, indel 1e-9, one Dirichlet
    signature; paired-end 150 bp reads (insert 350 +- 10), error rate 1e-3,
    normal sample included; coverage per job from batched_jobs."""
    lens, sig, specs = batched_jobs(seed=seed, n_jobs=n_jobs, **kw)
    g = core.Genome(ctx, lens, seed=seed + 1)
    engine = dict(snv_rate=[1e-8], indel_rate=[1e-9], sbs=sig[None, :], epoch_start=[0.0], exposure=[[1.0]],
                  n_pre_snv=0, n_pre_indel=0, pre_sbs=0, indel_max_len=10, indel_len_p=0.5, infinite_sites=True,
                  ploidy=2, drivers=[], germline=[])
    forests = []
    for s in specs:
        fd = core.synth_forest(growth=[0.2], death=[0.01], intro_at=[0], n_cells=2 * s["leaves"], n_samples=1,
                               sample_size=s["leaves"], seed=s["forest_seed"])
        forests.append(core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"],
                                   fd["n_samples"]))
    # every job's placement in one batched call
    t0 = time.perf_counter()
    muts = core.Mutations.place_batch(ctx, g, forests, [s["place_seed"] for s in specs], [engine] * len(specs))
    place_wall = (time.perf_counter() - t0) * 1e3
    place_ms = ctx.last_place_ms()
    jobs, reads = [], 0
    for s, f, m in zip(specs, forests, muts):
        seq = seq_job_params(s["coverage"])
        jobs.append(core.SeqJob(ctx, g, f, m, s["seq_seed"], **seq))
        reads += core.num_reads(g, f, **seq)
    return BatchedSet(g, jobs, specs, forests, muts, place_ms, reads, place_wall, engine)


@dataclass
class ServingPool:
    genome: core.Genome
    forests: list       # core.Forest
    muts: list          # core.Mutations
    n_samples: list
    cells: list
    descs: list = field(default_factory=list)  # the forests' node tables (synth_forest dicts)


def build_serving_pool(ctx: core.Context, seed: int = 555, n_forests: int = 8, total_bases: int = 50_000_000,
                       min_cells: int = 1_000, max_cells: int = 1_000_000) -> ServingPool:
    """Device-resident inputs of the serving workload: n_forests two-species
    forests with log-uniform cell counts in [min_cells, max_cells] (the
    smallest and largest included), 2-3 samples each of up to 1e4 cells, over
    a shared 50 Mb reference; SNV 1e-8, indel 1e-9."""
    rng = np.random.default_rng(seed)
    lens = log_uniform_lengths(4, total_bases, 1.0, 4.0, rng)
    g = core.Genome(ctx, lens, seed=seed + 1)
    sig = dirichlet_signature(rng)
    engine = dict(snv_rate=[1e-8, 1e-8], indel_rate=[1e-9, 1e-9], sbs=sig[None, :], epoch_start=[0.0],
                  exposure=[[1.0]], n_pre_snv=0, n_pre_indel=0, pre_sbs=0, indel_max_len=10, indel_len_p=0.5,
                  infinite_sites=True, ploidy=2, drivers=[], germline=[])
    sizes = np.exp(np.linspace(np.log(min_cells), np.log(max_cells), n_forests))
    forests, muts, ns, cells, descs = [], [], [], [], []
    for k, c in enumerate(sizes.astype(int)):
        n_s = 2 + k % 2
        ss = int(min(10_000, c // (2 * n_s)))
        fd = core.synth_forest(growth=[0.2, 0.3], death=[0.01, 0.01], intro_at=[0, max(2, c // 1000)], n_cells=int(c),
                               n_samples=n_s, sample_size=ss, seed=int(rng.integers(1 << 31)))
        f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
        m = core.Mutations.place(ctx, g, f, int(rng.integers(1 << 31)), **engine)
        forests.append(f)
        muts.append(m)
        ns.append(n_s)
        cells.append(int(c))
        descs.append(fd)
    return ServingPool(g, forests, muts, ns, cells, descs)


def serving_requests(pool: ServingPool, n: int, rate: float, seed: int = 9, min_cov: float = 30.0,
                     max_cov: float = 500.0):
    """n requests with Poisson arrivals at `rate` per second: (arrival s,
    forest index, seed, sequencing parameters); coverage log-uniform in
    [min_cov, max_cov], a random non-empty subset of the forest's samples,
    paired-end 150 bp reads, error rate 1e-3, normal sample included."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    out = []
    for i in range(n):
        k = int(rng.integers(len(pool.forests)))
        ns = pool.n_samples[k]
        pick = [s for s in range(ns) if rng.random() < 0.6] or [int(rng.integers(ns))]
        seq = seq_job_params(float(np.exp(rng.uniform(np.log(min_cov), np.log(max_cov)))))
        seq["samples"] = pick
        out.append((float(t[i]), k, int(rng.integers(1 << 31)), seq))
    return out


def cna_segments(lens, rng, n: int, species: int) -> list:
    """n allele-specific copy-number segments of one species: uniform start,
    log-uniform length in [1, 50] Mb (clipped to the chromosome), total copy
    state drawn from {0, 1, 3, 4} of a diploid segment: 0 deletes both
    lineages, 1 one lineage, 3 adds one copy to a lineage, 4 two copies."""
    from . import _abi as A
    out = []
    for _ in range(n):
        c = int(rng.integers(len(lens)))
        ln = int(lens[c])
        start = int(rng.integers(ln))
        length = int(min(np.exp(rng.uniform(np.log(1e6), np.log(50e6))), ln - start))
        state = int(rng.choice([0, 1, 3, 4]))
        a = int(rng.integers(2))
        if state == 0:
            out += [(species, c, start, length, 0, A.PSX_CNA_DEL, 0), (species, c, start, length, 1, A.PSX_CNA_DEL, 0)]
        elif state == 1:
            out.append((species, c, start, length, a, A.PSX_CNA_DEL, 0))
        else:
            out.append((species, c, start, length, a, A.PSX_CNA_AMP, state - 2))
    return out


def cna_purity(seed: int = 404, n_cells: int = 100_000, sample_size: int = 10_000,
               total_bases: int = 100_000_000) -> Workload:
    """Two species (growth 0.2; driver 0.3 introduced at 1,000 cells) grown to
    1e5 cells, one sample of 1e4 cells, a 100 Mb reference over 10
    chromosomes; eight copy-number segments in the founder and four in the
    driver clone (cna_segments), one whole-genome doubling on the driver
    clone; SNV 1e-8, indel 1e-9; 500x paired-end 150 bp reads (insert 350 +-
    10), error rate 1e-3, purity 0.6, normal sample included."""
    from . import _abi as A
    rng = np.random.default_rng(seed)
    lens = log_uniform_lengths(10, total_bases, 1.0, 5.0, rng)
    forest = core.synth_forest(growth=[0.2, 0.3], death=[0.01, 0.01], intro_at=[0, 1000], n_cells=n_cells,
                               n_samples=1, sample_size=sample_size, seed=seed + 1)
    sig = dirichlet_signature(rng)
    cna = cna_segments(lens, rng, 8, 0) + [(1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0)] + cna_segments(lens, rng, 4, 1)
    engine = dict(snv_rate=[1e-8, 1e-8], indel_rate=[1e-9, 1e-9], sbs=sig[None, :], epoch_start=[0.0],
                  exposure=[[1.0]], n_pre_snv=0, n_pre_indel=0, pre_sbs=0, indel_max_len=10, indel_len_p=0.5,
                  infinite_sites=True, ploidy=2, drivers=[], germline=[], cna=cna)
    seq = dict(coverage=500.0, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
               purity=0.6, with_normal_sample=True)
    return Workload("cna_purity", lens, seed + 2, forest, engine, seq, seed + 3, seed + 4,
                    notes=dict(n_cells=n_cells, sample_size=sample_size, genome=int(lens.sum()), chromosomes=10,
                               cna_drivers=len(cna)))
