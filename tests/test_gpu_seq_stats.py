"""The sequencing model's closed forms checked on the GPU's own outputs:
1,000 seeds of the single-end case of test_oracle_seq_stats run through
psx_simulate_seq (the count kernel), then the same tests (depth against the
exact sum of per-stratum binomials, alternative reads given depth against
Binomial(d, q), the VAF against its exact mixture; chi-squared and KS,
p > 0.01). A few seeds are also compared bit for bit with the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import test_oracle_seq_stats as T
from oracle import oracle as O
from process_amd import core

pytestmark = pytest.mark.gpu


def star_arrays(n_leaves: int):
    cid = np.arange(n_leaves + 1, dtype=np.int64)
    par = np.concatenate([[-1], np.zeros(n_leaves, np.int64)])
    sample = np.concatenate([[-1], np.zeros(n_leaves, np.int64)]).astype(np.int32)
    return cid, par, np.zeros(n_leaves + 1, np.uint32), sample, np.zeros(n_leaves + 1)


@pytest.fixture(scope="module")
def gpu_se_runs(ctx):
    c = T.SE_CASE
    G, L, C = c["G"], c["L"], c["C"]
    lens = np.array([G], np.uint32)
    cid, par, sp, sample, birth = star_arrays(c["n_leaves"])
    pf = core.Forest(ctx, cid, par, sp, sample, birth, 1)
    g = core.Genome(ctx, lens, seed=5)
    m = T.table(c["rows"])
    pm = core.Mutations.upload(ctx, pf, m)
    of = T.star_forest(c["n_leaves"])
    out = {}
    for e in (0.0, 0.01):
        alts, deps = [], []
        n = 0
        for s in T.SEEDS:
            a, d, p, n = core.simulate_seq_arrays(ctx, g, pf, pm, s, coverage=C, read_size=L, error_rate=e,
                                                  with_normal_sample=False)
            alts.append(a[0])
            deps.append(d[0])
            if s < 3:  # the GPU is the oracle, bit for bit
                oa, od, op, on = O.simulate_seq(lens, of, m, s, coverage=C, read_size=L, error_rate=e,
                                                with_normal_sample=False)
                assert np.array_equal(a, oa) and np.array_equal(d, od) and n == on
        out[e] = (np.array(alts), np.array(deps), n)
    pm.close()
    return dict(G=G, L=L, C=C, m=m, out=out, f=c["f"])


def test_gpu_read_count_exact(gpu_se_runs):
    r = gpu_se_runs
    assert r["out"][0.0][2] == round(r["C"] * r["G"] / r["L"])


def test_gpu_depth_distribution_matches_closed_form(gpu_se_runs):
    T.check_depth(gpu_se_runs)


@pytest.mark.parametrize("e", [0.0, 0.01])
def test_gpu_alt_given_depth_is_binomial(gpu_se_runs, e):
    T.check_alt_given_depth(gpu_se_runs, e)


def test_gpu_vaf_distribution_matches_mixture(gpu_se_runs):
    T.check_vaf(gpu_se_runs)
