"""The oracle pinned by what the mathematics fixes: closed-form moments of the
integer inversion tables, get_bin_dist's insert-size model, uniform synthetic
bases, the Euler-tour flattening against an ancestor walk, and invariants of
mutation placement (contexts match SBS classes, Poisson edge counts,
infinite-sites footprints, carrier sets)."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

import cases
from oracle import oracle as O
from process_amd import _abi as A


def table_pmf(T):
    T = np.asarray(T, np.float64)
    return np.diff(np.concatenate([[0.0], T])) / 2**32


# ------------------------------------------------------------------ tables

@pytest.mark.parametrize("lam", [0.02, 0.6, 6.0, 60.0, 600.0])
def test_poisson_table_moments(lam):
    k0, T = O.poisson_table(lam)
    p = table_pmf(T)
    k = k0 + np.arange(len(p))
    assert abs((p * k).sum() - lam) < 1e-6 * max(1.0, lam)
    assert abs((p * (k - lam) ** 2).sum() - lam) < 1e-5 * max(1.0, lam)


def test_binomial_table_moments():
    k0, T = O.binomial_table(490, 0.7142857142857143)
    p = table_pmf(T)
    k = k0 + np.arange(len(p))
    m = (p * k).sum()
    assert abs(m - 350.0) < 1e-6
    assert abs((p * (k - m) ** 2).sum() - 100.0) < 1e-4


def test_insert_size_model_of_get_bin_dist():
    """src/seq_simulation.cpp:431-451: q = sd^2/mean, p = 1 - q, t = (uint32)(mean/p);
    mean 350, sd 10 gives Binomial(490, 5/7) with mean 350 and variance 100."""
    q = 10 * 10 / 350
    p = 1 - q
    assert int(350 / p) == 490
    k0, T = O.binomial_table(490, p)
    pmf = table_pmf(T)
    k = k0 + np.arange(len(pmf))
    assert abs((pmf * k).sum() - 350) < 1e-6
    ref = stats.binom(490, p)
    assert np.allclose(pmf, ref.pmf(k), atol=2e-9)


def test_cdf_table_rejects_bad_input():
    with pytest.raises(O.OracleError):
        O.cdf_table([0.0, 0.0])
    with pytest.raises(O.OracleError):
        O.cdf_table([0.5, -0.1])


# ------------------------------------------------------------------ genome

def unpack(packed, lens):
    out, off = [], 0
    for ln in lens:
        nw = (int(ln) + 15) // 16
        w = packed[off: off + nw]
        b = ((w[:, None] >> (2 * np.arange(16, dtype=np.uint32))) & 3).reshape(-1)[: int(ln)]
        out.append(b.astype(np.uint8))
        off += nw
    return out


def test_genome_uniform_and_masked():
    lens = np.array([200_003, 100_000, 17], np.uint32)
    pk = O.genome_synth(lens, 9)
    assert len(pk) == sum((int(l) + 15) // 16 for l in lens)
    chrs = unpack(pk, lens)
    allb = np.concatenate(chrs)
    counts = np.bincount(allb, minlength=4)
    assert stats.chisquare(counts).pvalue > 1e-3
    # dinucleotides are independent
    di = np.bincount(chrs[0][:-1] * 4 + chrs[0][1:], minlength=16)
    assert stats.chisquare(di).pvalue > 1e-3
    # bits past a chromosome end are zero
    last = pk[(200_003 + 15) // 16 - 1]
    assert last >> (2 * (200_003 % 16)) == 0
    assert not np.array_equal(chrs[0][:1000], chrs[1][:1000])
    assert np.array_equal(O.genome_synth(lens, 9), pk)
    assert not np.array_equal(O.genome_synth(lens, 10), pk)


# ------------------------------------------------------------------ forest

def test_flatten_against_ancestor_walk():
    fd = cases.forest(n_cells=300, n_samples=3, sample_size=40, seed=5)
    f = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], 3)
    n = f.n_nodes
    assert n == len(fd["cell_id"])
    # DFS pre-order: parent before child, subtrees contiguous
    par = f.parent.astype(np.int64)
    assert all(par[v] < v for v in range(n) if par[v] != A.PSX_NO_NODE)
    is_leaf = np.ones(n, bool)
    is_leaf[par[par != A.PSX_NO_NODE]] = False
    rank = -np.ones(n, np.int64)
    rank[is_leaf] = np.arange(is_leaf.sum())
    assert f.n_leaves == is_leaf.sum()
    for v in range(0, n, 7):
        for leaf in np.nonzero(is_leaf)[0][::5]:
            u, anc = leaf, False
            while u != A.PSX_NO_NODE:
                if u == v:
                    anc = True
                    break
                u = par[u]
            assert anc == (f.leaf_lo[v] <= rank[leaf] < f.leaf_hi[v])
    # samples
    by_id = {int(c): i for i, c in enumerate(fd["cell_id"])}
    for s in range(3):
        want = sorted(rank[np.nonzero(np.isin(f.cell_id, fd["cell_id"][fd["sample"] == s]))[0]].tolist())
        got = f.sample_leaves[f.sample_begin[s]: f.sample_begin[s + 1]].tolist()
        assert got == want and len(got) == 40
    assert by_id


def test_flatten_rejects_malformed():
    with pytest.raises(O.OracleError, match="duplicated"):
        O.Forest([1, 1], [-1, 1], [0, 0], [-1, 0], [0, 0], 1)
    with pytest.raises(O.OracleError, match="unknown ancestor"):
        O.Forest([1, 2], [-1, 7], [0, 0], [-1, 0], [0, 0], 1)
    with pytest.raises(O.OracleError, match="cycle"):
        O.Forest([1, 2, 3], [-1, 3, 2], [0, 0, 0], [-1, -1, 0], [0, 0, 0], 1)
    with pytest.raises(O.OracleError, match="invalid sample"):
        O.Forest([1, 2], [-1, 1], [0, 0], [0, 0], [0, 0], 1)


# --------------------------------------------------------------- placement

@pytest.fixture(scope="module")
def placed():
    lens = np.array([120_000, 80_000], np.uint32)
    pk = O.genome_synth(lens, 21)
    fd = cases.forest(n_cells=800, n_samples=2, sample_size=200, seed=8)
    f = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], 2)
    eng = cases.engine(pk, lens, snv=4e-6, indel=4e-7, pre_snv=30, pre_indel=5)
    m = O.place_mutations(lens, pk, f, 99, **eng)
    return dict(lens=lens, pk=pk, fd=fd, f=f, eng=eng, m=m, bases=unpack(pk, lens))


COMP = np.array([3, 2, 1, 0])
SUB_TARGET = [0, 2, 3, 0, 1, 2]


def sbs_class(b5, ref, b3, alt):
    """COSMIC 96-class index of an SNV given its + strand context."""
    if ref in (0, 2):  # purine: use the reverse complement
        b5, ref, b3, alt = COMP[b3], COMP[ref], COMP[b5], COMP[alt]
    mid = 0 if ref == 1 else 1
    targets = SUB_TARGET[:3] if mid == 0 else SUB_TARGET[3:]
    sub = 3 * mid + targets.index(alt)
    return sub * 16 + b5 * 4 + b3


def test_snv_contexts_and_classes(placed):
    m, bases = placed["m"], placed["bases"]
    snv = m[(m["kind"] == A.PSX_SNV) & (m["nature"] == A.PSX_PASSENGER)]
    assert len(snv) > 300
    classes = []
    for r in snv:
        b = bases[r["chr"]]
        assert b[r["pos"]] == r["ref"] and r["alt"] != r["ref"]
        classes.append(sbs_class(int(b[r["pos"] - 1]), int(r["ref"]), int(b[r["pos"] + 1]), int(r["alt"])))
    classes = np.array(classes)
    # the class distribution follows the signature of the recorded cause
    eng = placed["eng"]
    for g in range(len(eng["sbs"])):
        sel = classes[snv["cause"] == g]
        if len(sel) < 200:
            continue
        # pool classes into 6 substitution types for a well-powered test
        obs = np.bincount(sel // 16, minlength=6)
        exp = eng["sbs"][g].reshape(6, 16).sum(axis=1) * len(sel)
        assert stats.chisquare(obs, exp).pvalue > 1e-4


def test_indels_and_references(placed):
    m, bases = placed["m"], placed["bases"]
    ind = m[m["kind"] != A.PSX_SNV]
    assert len(ind) > 20
    for r in ind:
        b = bases[r["chr"]]
        assert 1 <= r["len"] <= 10 and b[r["pos"]] == r["ref"]
        if r["kind"] == A.PSX_DEL:
            dele = [(int(r["seq"]) >> (2 * i)) & 3 for i in range(r["len"])]
            assert dele == b[r["pos"] + 1: r["pos"] + 1 + r["len"]].tolist()
    # geometric(1/2) lengths: P(1) = 1/2 (truncated at 10)
    frac1 = (ind["len"] == 1).mean()
    assert abs(frac1 - 0.5) < 4 * np.sqrt(0.25 / len(ind))


def test_infinite_sites_footprints_do_not_touch(placed):
    m = placed["m"]
    for c in range(len(placed["lens"])):
        r = m[m["chr"] == c]
        start = r["pos"].astype(np.int64)
        end = start + np.where(r["kind"] == A.PSX_DEL, r["len"], 0)
        o = np.argsort(start, kind="stable")
        start, end = start[o], end[o]
        assert np.all(start[1:] > np.maximum.accumulate(end)[:-1] + 1)


def test_poisson_edge_counts(placed):
    """Counts per non-root edge ~ Poisson(rate(species) * ploidy * G)."""
    m, f, eng, lens = placed["m"], placed["f"], placed["eng"], placed["lens"]
    G = int(lens.sum())
    pas = m[(m["nature"] == A.PSX_PASSENGER) & (m["kind"] == A.PSX_SNV)]
    cnt = np.bincount(pas["node"], minlength=f.n_nodes)
    species = np.asarray(placed["fd"]["species"])[np.argsort(placed["fd"]["cell_id"])]
    # species by DFS index via cell id
    sp_of_cell = dict(zip(placed["fd"]["cell_id"].tolist(), placed["fd"]["species"].tolist()))
    sp = np.array([sp_of_cell[int(c)] for c in f.cell_id])
    nonroot = f.parent != A.PSX_NO_NODE
    for s in (0, 1):
        sel = nonroot & (sp == s)
        if sel.sum() < 30:
            continue
        lam = eng["snv_rate"][s] * 2 * G
        x = cnt[sel]
        z = (x.sum() - lam * sel.sum()) / np.sqrt(lam * sel.sum())
        assert abs(z) < 4.5
        assert abs(x.var() / lam - 1) < 0.35
    assert species is not None
    assert cnt[~nonroot].sum() == 0


def test_carrier_sets_and_special_mutations(placed):
    m, f, eng = placed["m"], placed["f"], placed["eng"]
    pre = m[m["nature"] == A.PSX_PRENEOPLASTIC]
    assert len(pre) == eng["n_pre_snv"] + eng["n_pre_indel"]
    assert np.all(pre["leaf_lo"] == 0) and np.all(pre["leaf_hi"] == f.n_leaves)
    germ = m[m["nature"] == A.PSX_GERMINAL]
    assert len(germ) == len(eng["germline"])
    drv = m[m["nature"] == A.PSX_DRIVER]
    assert len(drv) >= 1
    par = f.parent
    sp_of_cell = dict(zip(placed["fd"]["cell_id"].tolist(), placed["fd"]["species"].tolist()))
    sp = np.array([sp_of_cell[int(c)] for c in f.cell_id])
    for r in drv:
        v = r["node"]
        assert sp[v] == 1 and (par[v] == A.PSX_NO_NODE or sp[par[v]] != 1)
    # passenger carrier range equals its node's subtree
    pas = m[m["nature"] == A.PSX_PASSENGER]
    assert np.array_equal(pas["leaf_lo"], f.leaf_lo[pas["node"]])
    assert np.array_equal(pas["leaf_hi"], f.leaf_hi[pas["node"]])


def test_sid_order_and_heads(placed):
    m = placed["m"]
    key = (m["chr"].astype(np.uint64) << 58) | (m["pos"].astype(np.uint64) << 26)
    assert np.all(np.diff(key.astype(np.float64)) >= 0)
    assert m["sid_head"][0] == 1
    assert np.array_equal(O.sort_table(m), m)


def test_placement_deterministic(placed):
    again = O.place_mutations(placed["lens"], placed["pk"], placed["f"], 99, **placed["eng"])
    assert again.tobytes() == placed["m"].tobytes()
    other = O.place_mutations(placed["lens"], placed["pk"], placed["f"], 100, **placed["eng"])
    assert other.tobytes() != placed["m"].tobytes()


def test_finite_sites_allows_collisions():
    lens = np.array([3_000], np.uint32)
    pk = O.genome_synth(lens, 2)
    fd = cases.forest(n_cells=300, n_samples=1, sample_size=100, seed=1)
    f = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], 1)
    eng = cases.engine(pk, lens, snv=3e-4, indel=0.0, pre_indel=0, infinite_sites=False)
    m = O.place_mutations(lens, pk, f, 5, **eng)
    assert len(np.unique(m["pos"])) < len(m)
    eng_is = dict(eng, infinite_sites=True, snv_rate=[1e-5, 1e-5])
    m2 = O.place_mutations(lens, pk, f, 5, **eng_is)
    assert len(np.unique(m2["pos"])) == len(m2)


def test_epoch_exposures():
    """Nodes born in an epoch draw signatures from that epoch's exposure."""
    lens = np.array([200_000], np.uint32)
    pk = O.genome_synth(lens, 4)
    fd = cases.forest(n_cells=600, n_samples=1, sample_size=300, seed=11)
    f = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], 1)
    birth = dict(zip(fd["cell_id"].tolist(), fd["birth"].tolist()))
    t_mid = float(np.median(fd["birth"]))
    sig = cases.signatures(2, 3)
    eng = dict(snv_rate=[5e-6, 5e-6], indel_rate=[0, 0], sbs=sig, epoch_start=[0.0, t_mid],
               exposure=[[1.0, 0.0], [0.0, 1.0]], n_pre_snv=0, n_pre_indel=0, infinite_sites=True)
    m = O.place_mutations(lens, pk, f, 3, **eng)
    node_birth = np.array([birth[int(c)] for c in f.cell_id])
    tb = node_birth[m["node"]]
    assert np.all(m["cause"][tb < t_mid] == 0)
    assert np.all(m["cause"][tb >= t_mid] == 1)
