"""The sequencing model pinned by closed forms over 1,000 seeds: depth equals
a sum of per-stratum binomials, alternative reads given depth are
Binomial(d, q) with q = f (1 - e) + (1 - f) e / 3, purity mixes in normal
genomes with probability 1 - purity, read counts are exact, and the VAF
distribution matches its exact mixture (chi-squared, p > 0.01)."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O
from process_amd import _abi as A

STRATUM = A.PSX_STRATUM
SEEDS = range(1000)


def star_forest(n_leaves: int, n_samples: int = 1):
    """A root with n_leaves children; leaf i goes to sample i % n_samples."""
    cid = np.arange(n_leaves + 1, dtype=np.int64)
    par = np.concatenate([[-1], np.zeros(n_leaves, np.int64)])
    sample = np.concatenate([[-1], np.arange(n_leaves) % n_samples]).astype(np.int32)
    return O.Forest(cid, par, np.zeros(n_leaves + 1, np.uint32), sample, np.zeros(n_leaves + 1), n_samples)


def table(rows):
    """rows: (chr, pos, leaf_lo, leaf_hi, allele, ref, alt, nature)."""
    m = np.zeros(len(rows), A.MUTATION_DTYPE)
    for i, (c, p, lo, hi, al, ref, alt, nat) in enumerate(rows):
        m[i] = (c, p, lo, hi, 0, 0, A.PSX_NO_CAUSE, al, A.PSX_SNV, ref, alt, 0, nat, 0, (0, 0, 0))
    return O.sort_table(m)


def depth_pmf(x, L, T, W, maxd=400):
    """Exact depth pmf at position x for single-end reads (sum over strata of
    Binomial(n_j, overlap_j / len_j))."""
    pmf = np.zeros(maxd)
    pmf[0] = 1.0
    J = (W + STRATUM - 1) // STRATUM
    for j in range(J):
        b, e = j * STRATUM, min((j + 1) * STRATUM, W)
        nj = T * e // W - T * b // W
        ov = max(0, min(e, x + 1) - max(b, x - L + 1))
        if ov == 0 or nj == 0:
            continue
        comp = stats.binom(nj, ov / (e - b)).pmf(np.arange(maxd))
        pmf = np.convolve(pmf, comp)[:maxd]
    return pmf


def chi2_discrete(obs_values, pmf, min_expected=5.0):
    """Chi-squared of observed integer values against a pmf, pooling sparse bins."""
    n = len(obs_values)
    counts = np.bincount(obs_values, minlength=len(pmf))[: len(pmf)]
    exp = pmf[: len(counts)] * n
    bins_o, bins_e, acc_o, acc_e = [], [], 0.0, 0.0
    for o, e in zip(counts, exp):
        acc_o += o
        acc_e += e
        if acc_e >= min_expected:
            bins_o.append(acc_o)
            bins_e.append(acc_e)
            acc_o = acc_e = 0.0
    if acc_e > 0 and bins_e:
        bins_o[-1] += acc_o
        bins_e[-1] += acc_e
    bins_e = np.array(bins_e)
    bins_e *= np.sum(bins_o) / bins_e.sum()
    return stats.chisquare(bins_o, bins_e).pvalue


# The single-end case: a 20 kb genome, 100 bp reads at 30x, a star forest of
# 10 leaves, three passengers at positions inside stratum 0, across the
# stratum 0/1 boundary and inside stratum 1.
SE_CASE = dict(G=20_000, L=100, C=30.0, n_leaves=10,
               rows=[(0, 5000, 0, 3, 0, 0, 1, A.PSX_PASSENGER),    # f = 3/10 * 1/2
                     (0, 8150, 0, 10, 1, 1, 2, A.PSX_PASSENGER),   # clonal, allele 1: f = 1/2
                     (0, 12000, 9, 10, 0, 2, 3, A.PSX_PASSENGER)],  # one cell: f = 1/20
               f=[0.15, 0.5, 0.05])


@pytest.fixture(scope="module")
def se_runs():
    """1,000 seeds of single-end 30x sequencing of a 20 kb genome."""
    G, L, C = SE_CASE["G"], SE_CASE["L"], SE_CASE["C"]
    lens = np.array([G], np.uint32)
    f = star_forest(SE_CASE["n_leaves"])
    m = table(SE_CASE["rows"])
    out = {}
    for e in (0.0, 0.01):
        alts, deps = [], []
        for s in SEEDS:
            a, d, p, n = O.simulate_seq(lens, f, m, s, coverage=C, read_size=L, error_rate=e,
                                        with_normal_sample=False)
            alts.append(a[0])
            deps.append(d[0])
        out[e] = (np.array(alts), np.array(deps), n)
    return dict(G=G, L=L, C=C, m=m, out=out, f=SE_CASE["f"])


def test_read_count_exact(se_runs):
    G, L, C = se_runs["G"], se_runs["L"], se_runs["C"]
    assert se_runs["out"][0.0][2] == round(C * G / L)


def test_depth_distribution_matches_closed_form(se_runs):
    check_depth(se_runs)


def check_depth(se_runs):
    G, L, C = se_runs["G"], se_runs["L"], se_runs["C"]
    T, W = round(C * G / L), G - L + 1
    alts, deps, _ = se_runs["out"][0.0]
    m = se_runs["m"]
    for k in range(len(m)):
        pmf = depth_pmf(int(m["pos"][k]), L, T, W)
        assert abs((pmf * np.arange(len(pmf))).sum() - C) < 0.5
        assert chi2_discrete(deps[:, k], pmf) > 0.01, f"position {m['pos'][k]}"
    # errors do not change depth
    assert np.array_equal(se_runs["out"][0.01][1], deps)


@pytest.mark.parametrize("e", [0.0, 0.01])
def test_alt_given_depth_is_binomial(se_runs, e):
    check_alt_given_depth(se_runs, e)


def check_alt_given_depth(se_runs, e):
    alts, deps, _ = se_runs["out"][e]
    m = se_runs["m"]
    fs = dict(zip(m["pos"].tolist(), se_runs["f"]))
    for k in range(len(m)):
        f = fs[int(m["pos"][k])]
        q = f * (1 - e) + (1 - f) * e / 3
        res = stats.binomtest(int(alts[:, k].sum()), int(deps[:, k].sum()), q)
        assert res.pvalue > 0.01
        # conditional distribution: PIT of alt given depth is uniform
        d, a = deps[:, k], alts[:, k]
        rng = np.random.default_rng(k)
        lo = np.where(a > 0, stats.binom.cdf(a - 1, d, q), 0.0)
        hi = stats.binom.cdf(a, d, q)
        pit = lo + rng.random(len(a)) * (hi - lo)
        assert stats.kstest(pit, "uniform").pvalue > 0.01


def test_vaf_distribution_matches_mixture(se_runs):
    check_vaf(se_runs)


def check_vaf(se_runs):
    G, L, C = se_runs["G"], se_runs["L"], se_runs["C"]
    T, W = round(C * G / L), G - L + 1
    alts, deps, _ = se_runs["out"][0.01]
    m = se_runs["m"]
    e = 0.01
    edges = np.linspace(0, 1, 11)
    for k, f in zip(range(len(m)), se_runs["f"]):
        pos = int(m["pos"][k])
        q = f * (1 - e) + (1 - f) * e / 3
        pmf_d = depth_pmf(pos, L, T, W)
        expv = np.zeros(len(edges) - 1)
        for d in range(1, len(pmf_d)):
            if pmf_d[d] < 1e-12:
                continue
            a = np.arange(d + 1)
            idx = np.minimum(np.searchsorted(edges, a / d, side="right") - 1, len(edges) - 2)
            np.add.at(expv, idx, pmf_d[d] * stats.binom.pmf(a, d, q))
        vaf = alts[:, k] / np.maximum(deps[:, k], 1)
        obs = np.bincount(np.minimum(np.searchsorted(edges, vaf, side="right") - 1, len(edges) - 2),
                          minlength=len(edges) - 1)
        keep = expv * len(vaf) >= 5
        o = np.append(obs[keep], obs[~keep].sum())
        x = np.append(expv[keep], expv[~keep].sum()) * len(vaf)
        if x[-1] < 5:
            o[-2] += o[-1]
            x[-2] += x[-1]
            o, x = o[:-1], x[:-1]
        x *= o.sum() / x.sum()
        assert stats.chisquare(o, x).pvalue > 0.01, f"position {pos}"


def test_purity_and_normal_sample():
    """Purity 0.6 on a clonal allele-0 passenger: f = 0.6 * 1/2; the normal
    sample carries germline (and, when asked, preneoplastic) mutations only."""
    G, L = 20_000, 150
    lens = np.array([G], np.uint32)
    f = star_forest(20)
    rows = [(0, 3000, 0, 20, 0, 0, 1, A.PSX_PASSENGER),
            (0, 9000, 0, 20, 1, 1, 0, A.PSX_GERMINAL),
            (0, 15000, 0, 20, 0, 3, 2, A.PSX_PRENEOPLASTIC)]
    m = table(rows)
    tot = {}
    for s in range(300):
        a, d, p, n = O.simulate_seq(lens, f, m, s, coverage=40, read_size=L, purity=0.6, with_normal_sample=True,
                                    preneoplastic_in_normal=False)
        for k in range(3):
            for so in range(2):
                ta, td = tot.get((so, k), (0, 0))
                tot[(so, k)] = (ta + int(a[so, k]), td + int(d[so, k]))
        assert p[1].tolist() == [0, 1, 0]  # normal sample holds the germline mutation only
        assert p[0].tolist() == [1, 1, 1]
    expect = {(0, 0): 0.3, (0, 1): 0.5, (0, 2): 0.5 * 0.6, (1, 0): 0.0, (1, 1): 0.5, (1, 2): 0.0}
    for key, q in expect.items():
        a, d = tot[key]
        if q == 0.0:
            assert a == 0
        else:
            assert stats.binomtest(a, d, q).pvalue > 0.01, key


def test_paired_end_depth_mean():
    G, L, C = 30_000, 150, 40.0
    lens = np.array([G], np.uint32)
    f = star_forest(4)
    m = table([(0, 10_000, 0, 4, 0, 0, 1, A.PSX_PASSENGER), (0, 20_000, 0, 4, 1, 0, 1, A.PSX_PASSENGER)])
    T = round(C * G / (2 * L))
    W = G - L + 1
    ds = []
    for s in range(400):
        a, d, p, n = O.simulate_seq(lens, f, m, s, coverage=C, read_size=L, insert_size_mean=350,
                                    insert_size_stddev=10, with_normal_sample=False)
        assert n == 2 * T
        ds.append(d[0])
    ds = np.array(ds, np.float64)
    mean = 2 * T * L / W  # mate 1 and mate 2 each cover a position with probability L / W
    assert abs(ds.mean() - mean) < 4 * np.sqrt(mean / ds.size) + 0.05


def test_include_non_sequenced_and_table_rows():
    lens = np.array([20_000], np.uint32)
    f = star_forest(6, n_samples=2)
    rows = [(0, 1000, 0, 1, 0, 0, 1, A.PSX_PASSENGER), (0, 1000, 2, 3, 0, 0, 1, A.PSX_PASSENGER),
            (0, 5000, 0, 6, 0, 1, 3, A.PSX_PASSENGER), (0, 19_990, 5, 6, 0, 2, 0, A.PSX_PASSENGER)]
    m = table(rows)
    assert m["sid_head"].tolist() == [1, 0, 1, 1]
    a, d, p, n = O.simulate_seq(lens, f, m, 3, coverage=5, read_size=100)
    r0, occ0, cov0, vaf0 = O.seq_table(m, a, d, p, include_non_sequenced=False)
    r1, occ1, cov1, vaf1 = O.seq_table(m, a, d, p, include_non_sequenced=True)
    assert set(r0.tolist()) <= set(r1.tolist())
    assert r1.tolist() == [0, 2, 3]  # heads only, one row per SID
    # duplicated SID: the head carries the counts of both records' carriers
    assert np.array_equal(cov1[:, 0], d[:, 0])
