"""The serving runtime on the GPU: requests of different sizes, parameters
and seeds submitted together under an HBM budget that holds only a few of
them at once give, each, exactly the tables of the same request run alone
(psx_simulate_seq), which the oracle also reproduces; the budget is never
exceeded, requests that cannot fit or are invalid are refused at submission,
and a loaded server keeps serving requests that arrive while launches run."""
from __future__ import annotations

import sys
import threading
import time

import numpy as np
import pytest

import cases
from oracle import oracle as O
from process_amd import core
from process_amd._lib import PsxError

pytestmark = pytest.mark.gpu

LENS = np.array([61_000, 37_500], np.uint32)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pool(ctx):
    say("[serve] building 3 forests over a 98.5 kb genome")
    packed = O.genome_synth(LENS, 41)
    g = core.Genome(ctx, LENS, seed=41)
    items = []
    for k, (cells, ns, ss) in enumerate([(300, 2, 60), (900, 3, 150), (2000, 1, 700)]):
        fd = cases.forest(n_cells=cells, n_samples=ns, sample_size=ss, seed=10 + k)
        f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
        eng = cases.engine(packed, LENS, snv=3e-6 * (k + 1), indel=3e-7)
        m = core.Mutations.place(ctx, g, f, 500 + k, **eng)
        of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
        items.append((f, m, of))
    return g, items


def requests():
    base = [dict(coverage=30, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3),
            dict(coverage=12, read_size=100, error_rate=5e-3, purity=0.7),
            dict(coverage=45, read_size=150, insert_size_mean=300, insert_size_stddev=12, error_rate=1e-3,
                 with_normal_sample=False),
            dict(coverage=20, read_size=90, chromosomes=[1], samples=[0], error_rate=2e-3)]
    out = []
    for i in range(14):
        seq = dict(base[i % len(base)])
        if "samples" in seq:
            seq["samples"] = [0]
        out.append((i % 3, 1000 + 7 * i, seq))
    return out


def test_results_identical_alone_and_in_the_mix(ctx, pool):
    g, items = pool
    reqs = requests()
    say(f"[serve] {len(reqs)} requests alone")
    alone = []
    for fi, seed, seq in reqs:
        f, m, _ = items[fi]
        alone.append(core.simulate_seq_arrays(ctx, g, f, m, seed, **seq))
    # the oracle agrees with two of them
    for k in (0, 5):
        fi, seed, seq = reqs[k]
        f, m, of = items[fi]
        oa, od, op, on = O.simulate_seq(LENS, of, m.download(), seed, **seq)
        assert np.array_equal(alone[k][0], oa) and np.array_equal(alone[k][1], od) and alone[k][3] == on
    sizes = [a[0].size * 9 for a in alone]
    budget = max(1 << 20, 3 * max(sizes))  # a few requests at a time
    say(f"[serve] mix: budget {budget} bytes, outputs {min(sizes)}..{max(sizes)} bytes, 2 lanes")
    srv = core.Server(ctx, budget, max_inflight=2, max_batch=4)
    try:
        tickets = [srv.submit(g, items[fi][0], items[fi][1], seed, **seq) for fi, seed, seq in reqs]
        for k, t in enumerate(tickets):
            a, d, p, st = srv.wait(t, timeout_ms=60_000)
            assert np.array_equal(a, alone[k][0]), k
            assert np.array_equal(d, alone[k][1]), k
            assert np.array_equal(p, alone[k][2]), k
            assert st["n_reads"] == alone[k][3]
            assert 1 <= st["batch_size"] <= 4 and st["latency_ms"] >= st["queued_ms"] >= 0
        tot = srv.totals()
        say(f"[serve] totals {tot}")
        assert tot["completed"] == len(reqs) and tot["failed"] == 0
        assert tot["peak_bytes"] <= tot["budget_bytes"] <= budget
        assert tot["launches"] >= 2
    finally:
        srv.close()


def test_admission_refusals(ctx, pool):
    g, items = pool
    f, m, _ = items[2]
    srv = core.Server(ctx, 1 << 20, max_inflight=1, max_batch=2)
    try:
        big = dict(coverage=5, read_size=100)
        need = core.num_out_samples(f, **big) * m.n * 9
        if need > (1 << 20):
            with pytest.raises(PsxError, match="serving budget"):
                srv.submit(g, f, m, 1, **big)
        with pytest.raises(PsxError, match="purity"):
            srv.submit(g, items[0][0], items[0][1], 1, coverage=5, read_size=100, purity=1.5)
        t = srv.submit(g, items[0][0], items[0][1], 3, coverage=5, read_size=100)
        a, d, p, st = srv.wait(t, timeout_ms=60_000)
        ref = core.simulate_seq_arrays(ctx, g, items[0][0], items[0][1], 3, coverage=5, read_size=100)
        assert np.array_equal(a, ref[0]) and np.array_equal(d, ref[1])
    finally:
        srv.close()


@pytest.mark.parametrize("overlap_us", [0, 250, 100_000])
def test_arrivals_while_running(ctx, pool, overlap_us):
    """A producer thread submits at intervals while the lanes run; every
    request completes and matches its lone result, whether a second lane
    starts at once, after a running launch's first 250 us or (100 ms) only
    when max_batch requests wait."""
    g, items = pool
    f, m, _ = items[1]
    seq = dict(coverage=25, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3)
    ref = {s: core.simulate_seq_arrays(ctx, g, f, m, s, **seq) for s in range(20, 30)}
    srv = core.Server(ctx, 64 << 20, max_inflight=2, max_batch=8, overlap_after_us=overlap_us)
    tickets = {}
    try:
        def produce():
            for s in range(20, 30):
                tickets[s] = srv.submit(g, f, m, s, **seq)
                time.sleep(0.002)

        th = threading.Thread(target=produce)
        th.start()
        th.join(timeout=60)
        assert not th.is_alive()
        for s, t in tickets.items():
            a, d, p, st = srv.wait(t, timeout_ms=60_000)
            assert np.array_equal(a, ref[s][0]) and np.array_equal(d, ref[s][1]) and np.array_equal(p, ref[s][2])
        assert srv.totals()["completed"] == 10
    finally:
        srv.close()


def test_copy_number_requests_in_the_mix(ctx, pool):
    """Requests on a table with copy-number events (amplification, deletion,
    whole-genome doubling, purity < 1) served together with plain ones: each
    result, and its count of accepted reads, equals the request run alone,
    and the oracle reproduces one of them."""
    from process_amd import _abi as A

    g, items = pool
    packed = O.genome_synth(LENS, 41)
    fd = core.synth_forest(growth=[0.2, 0.45], death=[0.01, 0.01], intro_at=[0, 20], n_cells=600, n_samples=2,
                           sample_size=90, seed=17)
    f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    eng = cases.engine(packed, LENS, snv=5e-6, indel=5e-7, pre_snv=10, pre_indel=2)
    eng["cna"] = [(0, 0, 4_000, 20_000, 0, A.PSX_CNA_AMP, 2), (0, 1, 2_000, 9_000, 1, A.PSX_CNA_DEL, 0),
                  (1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0), (1, 0, 30_000, 12_000, 1, A.PSX_CNA_AMP, 1)]
    m = core.Mutations.place(ctx, g, f, 77, **eng)
    assert len(m.cna()) > 0
    cna_seq = [dict(coverage=30, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3,
                    purity=0.6),
               dict(coverage=20, read_size=100, error_rate=2e-3)]
    reqs = [(f, m, 300 + k, cna_seq[k % 2]) for k in range(4)]
    reqs += [(items[k % 3][0], items[k % 3][1], 400 + k, dict(coverage=15, read_size=100, error_rate=1e-3))
             for k in range(4)]
    say(f"[serve] {len(reqs)} requests, 4 with copy-number events, alone")
    alone = [core.simulate_seq_arrays(ctx, g, rf, rm, seed, **seq) for rf, rm, seed, seq in reqs]
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    oa, od, op, on = O.simulate_seq(LENS, of, m.download(), 300, cna=m.cna(), **cna_seq[0])
    assert np.array_equal(alone[0][0], oa) and np.array_equal(alone[0][1], od) and alone[0][3] == on
    srv = core.Server(ctx, 64 << 20, max_inflight=2, max_batch=8)
    try:
        tickets = [srv.submit(g, rf, rm, seed, **seq) for rf, rm, seed, seq in reqs]
        for k, t in enumerate(tickets):
            a, d, p, st = srv.wait(t, timeout_ms=60_000)
            assert np.array_equal(a, alone[k][0]) and np.array_equal(d, alone[k][1]), k
            assert np.array_equal(p, alone[k][2]), k
            assert st["n_reads"] == alone[k][3], k
        assert srv.totals()["completed"] == len(reqs)
    finally:
        srv.close()


def test_freshly_uploaded_tables(ctx, pool):
    """Replayed (uploaded) tables are indexed for their largest positions, so
    the first request on a real genome builds the genome's stratum index while
    the serving lanes run on their own streams: every such request, submitted
    before any other call touched its table, equals the oracle, and the same
    table indexed for two genomes of different lengths in one batch keeps
    both indexes valid."""
    g, items = pool
    seq = dict(coverage=20, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3)
    recs = [items[k][1].download() for k in range(3)]
    ups = [core.Mutations.upload(ctx, items[k][0], recs[k]) for k in range(3)]
    lens2 = np.array([LENS[0], LENS[1] + 2_000], np.uint32)  # the same bases up to LENS, a longer chromosome 1
    g2 = core.Genome(ctx, lens2, seed=41)
    srv = core.Server(ctx, 64 << 20, max_inflight=2, max_batch=8)
    try:
        tickets = [srv.submit(g, items[k][0], ups[k], 60 + k, **seq) for k in range(3)]
        tickets.append(srv.submit(g2, items[0][0], ups[0], 70, **seq))
        got = [srv.wait(t, timeout_ms=60_000) for t in tickets]
    finally:
        srv.close()
    for k in range(3):
        oa, od, op, on = O.simulate_seq(LENS, items[k][2], recs[k], 60 + k, **seq)
        assert np.array_equal(got[k][0], oa) and np.array_equal(got[k][1], od), k
        assert np.array_equal(got[k][2], op) and got[k][3]["n_reads"] == on, k
    oa, od, op, on = O.simulate_seq(lens2, items[0][2], recs[0], 70, **seq)
    assert np.array_equal(got[3][0], oa) and np.array_equal(got[3][1], od) and got[3][3]["n_reads"] == on
    # one batch, one table, two genomes
    jobs = [core.SeqJob(ctx, g, items[0][0], ups[0], 80, **seq), core.SeqJob(ctx, g2, items[0][0], ups[0], 81, **seq)]
    core.simulate_seq_batch(ctx, jobs)
    for j, lens in zip(jobs, (LENS, lens2)):
        oa, od, op, on = O.simulate_seq(lens, items[0][2], recs[0], j.seed, **seq)
        a, d, p = j.results()
        assert np.array_equal(a, oa) and np.array_equal(d, od) and np.array_equal(p, op)
