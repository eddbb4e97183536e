"""Copy numbers in the oracle: placement of copy-number drivers and the
sequencing model with allele-specific copy numbers and normal admixture,
pinned by closed forms (expected depth and VAF from the copy-number profile
of the sample's cells), exact invariants (a deleted lineage yields no
alternative reads; without events the plain model is reproduced draw for
draw) and the event bookkeeping of the forest."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core

LENS = np.array([90_000], np.uint32)
AMP_REGION = (10_000, 40_000)   # allele 0 gains 2 copies in the founder
DEL_REGION = (50_000, 70_000)   # allele 1 lost in the founder


def cna_drivers(wgd=True):
    d = [(0, 0, AMP_REGION[0], AMP_REGION[1] - AMP_REGION[0], 0, A.PSX_CNA_AMP, 2),
         (0, 0, DEL_REGION[0], DEL_REGION[1] - DEL_REGION[0], 1, A.PSX_CNA_DEL, 0)]
    if wgd:
        d.append((1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0))
    return d


@pytest.fixture(scope="module")
def world():
    packed = O.genome_synth(LENS, 9)
    # the doubling clone (species 1) holds about half of the sampled cells
    fd = core.synth_forest(growth=[0.2, 0.45], death=[0.01, 0.01], intro_at=[0, 20], n_cells=500, n_samples=1,
                           sample_size=150, seed=3)
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    eng = cases.engine(packed, LENS, snv=4e-6, indel=0.0, pre_snv=30, pre_indel=0, drivers=False, germline=False)
    eng["cna"] = cna_drivers()
    recs = O.place_mutations(LENS, packed, of, 77, **eng)
    ev = O.place_cna(LENS, of, **eng)
    return dict(packed=packed, fd=fd, of=of, eng=eng, recs=recs, ev=ev)


def profile(ev, leaf, x, ploidy=2):
    """Copy numbers per allele lineage: one each, then the leaf's events in order."""
    cn = [1] * ploidy
    for e in ev:
        if not (e["leaf_lo"] <= leaf < e["leaf_hi"]):
            continue
        if e["kind"] == A.PSX_CNA_WGD:
            cn = [2 * c for c in cn]
        elif e["start"] <= x < e["end"]:
            cn[e["allele"]] = cn[e["allele"]] + e["copies"] if e["kind"] == A.PSX_CNA_AMP else 0
    return cn


def test_events_on_first_appearance_nodes(world):
    of, fd, ev = world["of"], world["fd"], world["ev"]
    leaves = of.sample_leaves
    wgd = ev[ev["kind"] == A.PSX_CNA_WGD]
    doubled = np.mean([any(e["leaf_lo"] <= l < e["leaf_hi"] for e in wgd) for l in leaves])
    assert 0.3 < doubled < 0.7
    species = {int(c): int(s) for c, s in zip(fd["cell_id"], fd["species"])}
    sp = np.array([species[int(c)] for c in of.cell_id])
    first = [v for v in range(of.n_nodes) if of.parent[v] == A.PSX_NO_NODE or sp[of.parent[v]] != sp[v]]
    want = sum(2 if sp[v] == 0 else 1 for v in first)
    assert len(ev) == want
    assert (np.diff(ev["node"].astype(np.int64)) >= 0).all()  # DFS order: ancestors first
    for e in ev:
        v = int(e["node"])
        assert v in first and (e["leaf_lo"], e["leaf_hi"]) == (of.leaf_lo[v], of.leaf_hi[v])
        if e["kind"] == A.PSX_CNA_WGD:
            assert sp[v] == 1 and e["chr"] == A.PSX_ALL_CHROMOSOMES
        else:
            assert sp[v] == 0 and e["chr"] == 0


def test_no_events_is_the_plain_model(world):
    seq = dict(coverage=20, read_size=100, insert_size_mean=300, insert_size_stddev=10, error_rate=1e-3, purity=0.8)
    a = O.simulate_seq(LENS, world["of"], world["recs"], 5, **seq)
    b = O.simulate_seq(LENS, world["of"], world["recs"], 5, cna=np.zeros(0, A.CNA_EVENT_DTYPE), **seq)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def expected_fraction(ev, leaves, x, purity, ploidy=2):
    """Tumour copies and carried-allele copies averaged over the sample's cells."""
    tot = np.array([sum(profile(ev, l, x)) for l in leaves], float)
    a0 = np.array([profile(ev, l, x)[0] for l in leaves], float)
    a1 = np.array([profile(ev, l, x)[1] for l in leaves], float)
    dna = purity * tot.mean() + (1 - purity) * ploidy
    return dna / ploidy, a0.mean() * purity / dna, a1.mean() * purity / dna


@pytest.mark.parametrize("purity", [1.0, 0.6])
def test_depth_and_vaf_follow_copy_numbers(world, purity):
    """Depth at x is coverage * (purity E[C] + (1 - purity) ploidy) / ploidy;
    a truncal SNV on lineage a has VAF purity E[c_a] / (that DNA quantity)."""
    of, recs, ev = world["of"], world["recs"], world["ev"]
    leaves = of.sample_leaves[of.sample_begin[0]:of.sample_begin[1]]
    pre = np.where((recs["nature"] == A.PSX_PRENEOPLASTIC) & (recs["kind"] == A.PSX_SNV))[0]
    assert len(pre) >= 20
    seq = dict(coverage=60, read_size=100, purity=purity, with_normal_sample=False)
    dep, alt = [], []
    for seed in range(12):
        a, d, p, n = O.simulate_seq(LENS, of, recs, 100 + seed, cna=ev, **seq)
        dep.append(d[0, pre])
        alt.append(a[0, pre])
    dep = np.array(dep, float)
    alt = np.array(alt, float)
    for k, q in enumerate(pre):
        x = int(recs["pos"][q])
        if x < 200 or x > LENS[0] - 200:
            continue
        # away from CN boundaries only
        if any(abs(x - b) < 150 for b in (*AMP_REGION, *DEL_REGION)):
            continue
        scale, f0, f1 = expected_fraction(ev, leaves, x, purity)
        d_exp = 60 * scale
        assert abs(dep[:, k].mean() - d_exp) < 5 * np.sqrt(d_exp / 12) + 1, (x, dep[:, k].mean(), d_exp)
        f = f0 if recs["allele"][q] == 0 else f1
        vaf = alt[:, k].sum() / dep[:, k].sum()
        sd = np.sqrt(max(f * (1 - f), 1e-3) / dep[:, k].sum())
        assert abs(vaf - f) < 5 * sd + 0.02, (x, vaf, f)


def test_deleted_lineage_has_no_alternative_reads(world):
    """Purity 1: an SNV on lineage 1 inside the founder's deletion of lineage
    1 is read from no cell, so it collects no alternative read (errors off)."""
    of, recs, ev = world["of"], world["recs"], world["ev"]
    a, d, p, n = O.simulate_seq(LENS, of, recs, 3, cna=ev, coverage=80, read_size=100, with_normal_sample=False)
    inside = (recs["pos"] >= DEL_REGION[0]) & (recs["pos"] < DEL_REGION[1]) & (recs["allele"] == 1) & \
        (recs["kind"] == A.PSX_SNV)
    assert inside.any()
    assert (a[0, inside] == 0).all()
    assert (d[0, inside] > 0).all()  # lineage 0 is still read there


def test_read_count_is_the_accepted_templates(world):
    """n_reads counts accepted reads: its mean matches the DNA quantity of the
    sample integrated over the chromosome."""
    of, recs, ev = world["of"], world["recs"], world["ev"]
    leaves = of.sample_leaves[of.sample_begin[0]:of.sample_begin[1]]
    seq = dict(coverage=30, read_size=100, purity=0.7, with_normal_sample=False)
    ns = [O.simulate_seq(LENS, of, recs, s, cna=ev, **seq)[3] for s in range(6)]
    xs = np.arange(50, LENS[0] - 100, 500)
    scale = np.mean([expected_fraction(ev, leaves, int(x), 0.7)[0] for x in xs])
    expect = 30 * LENS[0] / 100 * scale
    assert abs(np.mean(ns) - expect) < 0.02 * expect


def test_invalid_copy_number_driver_rejected():
    fd = cases.forest(n_cells=60, n_samples=1, sample_size=20, seed=2)
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    packed = O.genome_synth(LENS, 1)
    eng = cases.engine(packed, LENS, snv=0.0, indel=0.0, pre_snv=0, pre_indel=0, drivers=False, germline=False)
    eng["cna"] = [(0, 0, 10, 100, 3, A.PSX_CNA_AMP, 1)]  # allele 3 of a diploid genome
    with pytest.raises(O.OracleError):
        O.place_cna(LENS, of, **eng)


# ---------------------------------------------------------- passenger CNAs

CATALOGUE = [(0, 1_000, 5_000, A.PSX_CNA_AMP), (0, 20_000, 3_000, A.PSX_CNA_DEL),
             (0, 60_000, 10_000, A.PSX_CNA_AMP), (0, 80_000, 2_500, A.PSX_CNA_DEL)]


def passenger_engine(world, rates=(2e-6, 6e-6), catalogue=CATALOGUE):
    eng = dict(world["eng"])
    eng["cna"] = cna_drivers(wgd=False)
    eng["cna_rate"] = list(rates)
    eng["cna_catalogue"] = catalogue
    return eng


def test_passenger_cna_counts_are_poisson(world):
    """Per non-root edge Poisson(cna_rate(species) x ploidy x G) passenger
    events: over 300 placements, the total's mean and variance equal the
    closed-form sum of the edges' means (rates chosen so it is ~ 60)."""
    of, fd = world["of"], world["fd"]
    eng = passenger_engine(world, rates=(1e-7, 3e-7))
    flat_parent = of.parent
    G = int(LENS.sum())
    lam = sum((1e-7 if of.species[v] == 0 else 3e-7) * 2 * G for v in range(of.n_nodes)
              if flat_parent[v] != A.PSX_NO_NODE)
    n_drv = len(O.place_cna(LENS, of, seed=0, **dict(eng, cna_rate=None)))
    totals = np.array([len(O.place_cna(LENS, of, seed=s, **eng)) - n_drv for s in range(300)], np.float64)
    se = np.sqrt(lam / 300)
    assert abs(totals.mean() - lam) < 4 * se, (totals.mean(), lam)
    # variance of a Poisson total equals its mean (chi-square on 299 degrees of freedom, p in (0.001, 0.999))
    from scipy import stats
    chi = totals.var(ddof=1) * 299 / lam
    assert stats.chi2.ppf(0.001, 299) < chi < stats.chi2.ppf(0.999, 299)


def test_passenger_cna_events(world):
    """Passenger events sit on non-root edges after that node's drivers, in
    DFS order; regions are uniform over the catalogue and alleles uniform
    (chi-square over 200 placements); amplifications add one copy."""
    from scipy import stats
    of = world["of"]
    eng = passenger_engine(world, rates=(5e-7, 5e-7))
    drv = O.place_cna(LENS, of, seed=1, **dict(eng, cna_rate=None))
    region, allele = np.zeros(4), np.zeros(2)
    for s in range(200):
        ev = O.place_cna(LENS, of, seed=s, **eng)
        assert (np.diff(ev["node"].astype(np.int64)) >= 0).all()
        is_drv = np.zeros(len(ev), bool)
        k = 0
        for i, e in enumerate(ev):  # the drivers keep their order and precede the node's passengers
            if k < len(drv) and e.tobytes() == drv[k].tobytes():
                is_drv[i] = True
                k += 1
        assert k == len(drv)
        pas = ev[~is_drv]
        assert (of.parent[pas["node"]] != A.PSX_NO_NODE).all()
        for e in pas:
            key = (int(e["chr"]), int(e["start"]), int(e["end"] - e["start"]), int(e["kind"]))
            j = [(c, st, ln, kd) for c, st, ln, kd in CATALOGUE].index(key)
            region[j] += 1
            allele[e["allele"]] += 1
            assert e["copies"] == (1 if e["kind"] == A.PSX_CNA_AMP else 0)
            assert e["leaf_lo"] == of.leaf_lo[e["node"]] and e["leaf_hi"] == of.leaf_hi[e["node"]]
        # a node's passengers come after its drivers
        for v in set(drv["node"].tolist()):
            idx = np.nonzero(ev["node"] == v)[0]
            assert is_drv[idx[: (drv["node"] == v).sum()]].all()
    assert region.sum() > 500
    assert stats.chisquare(region).pvalue > 0.001 and stats.chisquare(allele).pvalue > 0.001


def test_passenger_cna_errors(world):
    of = world["of"]
    with pytest.raises(RuntimeError, match="need a CNA catalogue"):
        O.place_cna(LENS, of, seed=1, **passenger_engine(world, catalogue=[]))
    with pytest.raises(RuntimeError, match="invalid passenger CNA catalogue region 0"):
        O.place_cna(LENS, of, seed=1, **passenger_engine(world, catalogue=[(0, 89_000, 5_000, A.PSX_CNA_AMP)]))
    with pytest.raises(RuntimeError, match="CNA rates must be non-negative"):
        O.place_cna(LENS, of, seed=1, **passenger_engine(world, rates=(-1.0, 0.0)))


def test_random_allele_cna_drivers(world):
    """A CNA driver with no allele draws one per placement node, uniformly."""
    of = world["of"]
    eng = dict(world["eng"])
    eng["cna"] = [(0, 0, 1_000, 2_000, A.PSX_CNA_RANDOM_ALLELE, A.PSX_CNA_DEL, 0)]
    seen = set()
    for s in range(40):
        ev = O.place_cna(LENS, of, seed=s, **eng)
        assert len(ev) >= 1
        seen |= set(ev["allele"].tolist())
    assert seen == {0, 1}
