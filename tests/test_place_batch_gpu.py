"""Batched placement (psx_place_mutations_batch): many forests over one
genome in one call, each table byte for byte the single-forest placement's
and the oracle's, with the infinite-sites model on and off, ID signatures,
copy-number drivers, a coarse context sampling, an empty table, jobs whose
redraw rounds outlast the others' and many tiny jobs sharing waves; a bad job
fails the batch with its index in the message."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core
from process_amd._lib import PsxError

pytestmark = pytest.mark.gpu

LENS = [400_000, 150_000, 60_000]


def forests(ctx, specs):
    out = []
    for n_cells, seed, two in specs:
        fd = cases.forest(n_cells=n_cells, n_samples=2, sample_size=min(100, n_cells // 4), seed=seed,
                          two_species=two)
        f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
        of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
        out.append((f, of, two))
    return out


def engines(packed, fs):
    rng = np.random.default_rng(4)
    ids = rng.dirichlet(np.ones(A.PSX_ID_CLASSES), size=2)
    e = []
    for k, (_, _, two) in enumerate(fs):
        if two:
            eng = cases.engine(packed, LENS, snv=(1 + k) * 4e-7, indel=1e-7, pre_snv=3 * k, pre_indel=k,
                               infinite_sites=k != 2)
        else:
            sig = cases.signatures(1, k)
            eng = dict(snv_rate=[6e-7], indel_rate=[5e-8], sbs=sig, epoch_start=[0.0], exposure=[[1.0]],
                       n_pre_snv=2, n_pre_indel=0, pre_sbs=0, indel_max_len=10, indel_len_p=0.5,
                       infinite_sites=True, ploidy=2, drivers=[], germline=[])
        e.append(eng)
    e[1].update(id_sig=ids, id_exposure=[[0.3, 0.7], [0.6, 0.4]], pre_id=1)
    e[3]["context_sampling"] = 20
    e[4]["cna"] = [(1, 0, 1000, 20_000, 0, A.PSX_CNA_AMP, 2), (1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0)]
    # an empty table: no rates, no drivers, no germline, no preneoplastic mutations
    e[5].update(snv_rate=[0.0], indel_rate=[0.0], n_pre_snv=0)
    # dense: about one mutation per 60 bases, many more redraw rounds than the other jobs
    e[6].update(snv_rate=[3e-5, 3e-5], indel_rate=[3e-6, 3e-6])
    return e


def test_batch_equals_single_and_oracle(ctx):
    g = core.Genome(ctx, LENS, seed=13)
    packed = g.download()
    specs = [(800, 1, True), (1200, 2, True), (600, 3, True), (900, 4, True), (700, 5, True), (300, 6, False),
             (1000, 7, True)]
    fs = forests(ctx, specs)
    es = engines(packed, fs)
    seeds = [101 + 7 * k for k in range(len(fs))]
    batch = core.Mutations.place_batch(ctx, g, [f for f, _, _ in fs], seeds, es)
    assert len(batch) == len(fs)
    for k, ((f, of, _), eng, sd) in enumerate(zip(fs, es, seeds)):
        got = batch[k].download()
        one = core.Mutations.place(ctx, g, f, sd, **eng)
        assert got.tobytes() == one.download().tobytes(), k
        assert np.array_equal(batch[k].cna(), one.cna()), k
        assert got.tobytes() == O.place_mutations(LENS, packed, of, sd, **eng).tobytes(), k
    assert batch[5].n == 0
    assert batch[6].n > 5 * batch[0].n
    # the tables share one block: freeing some leaves the others whole
    keep = batch[6].download()
    for k in range(6):
        batch[k].close()
    assert batch[6].download().tobytes() == keep.tobytes()


def test_batch_error_names_the_job(ctx):
    g = core.Genome(ctx, LENS, seed=13)
    packed = g.download()
    fs = forests(ctx, [(400, 1, True), (400, 2, True), (400, 3, True)])
    es = [cases.engine(packed, LENS) for _ in fs]
    pos = es[2]["drivers"][0][2]
    base = cases.base_of(packed, LENS, 0, pos)
    es[2]["drivers"] = [(1, 0, pos, base, -1)]  # alternative equal to the reference base
    with pytest.raises(PsxError, match="job 2: invalid driver 0"):
        core.Mutations.place_batch(ctx, g, [f for f, _, _ in fs], [1, 2, 3], es)
    es[2] = cases.engine(packed, LENS)
    es[0]["ploidy"] = 0
    with pytest.raises(PsxError, match="job 0: invalid ploidy"):
        core.Mutations.place_batch(ctx, g, [f for f, _, _ in fs], [1, 2, 3], es)


def test_many_tiny_jobs_share_waves(ctx):
    """48 forests of 12-60 cells: each job has a few to tens of mutations and
    nodes, so most waves of every per-node and per-mutation kernel hold
    several jobs (one pass per job in the wave); every table equals the
    single-forest placement, and every eighth the oracle's."""
    g = core.Genome(ctx, LENS, seed=21)
    packed = g.download()
    rng = np.random.default_rng(9)
    fs, es, seeds = [], [], []
    for k in range(48):
        n = int(rng.integers(12, 61))
        fd = cases.forest(n_cells=n, n_samples=1, sample_size=max(2, n // 4), seed=100 + k)
        f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
        of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
        fs.append((f, of))
        es.append(cases.engine(packed, LENS, snv=float(rng.uniform(2e-6, 2e-5)), indel=2e-7, drivers=k % 3 == 0,
                               germline=k % 4 == 0, infinite_sites=k % 5 != 0))
        seeds.append(500 + k)
    batch = core.Mutations.place_batch(ctx, g, [f for f, _ in fs], seeds, es)
    sizes = []
    for k, ((f, of), eng, sd) in enumerate(zip(fs, es, seeds)):
        got = batch[k].download()
        sizes.append(len(got))
        assert got.tobytes() == core.Mutations.place(ctx, g, f, sd, **eng).download().tobytes(), k
        if k % 8 == 0:
            assert got.tobytes() == O.place_mutations(LENS, packed, of, sd, **eng).tobytes(), k
    assert sum(sizes) > 48 * 5 and min(sizes) < 64
