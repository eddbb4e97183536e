"""Indel placement from COSMIC ID83 signatures in the oracle: the reference's
ID classes (sites per class, the site lists of sparse classes, the site test)
against a brute-force enumeration over the bases as strings, against closed
forms on i.i.d. uniform references, and placement (class histogram against
the signature renormalised over the classes with a site; every placed indel
sits on a site of its class and deletes or inserts its unit after the anchor
base)."""
from __future__ import annotations

import os
from collections import defaultdict

import numpy as np
import pytest
from scipy import stats

import cases
from oracle import oracle as O
from process_amd import _abi as A

BASE = "ACGT"
NC = A.PSX_ID_CLASSES


def pack(seqs):
    words = []
    for s in seqs:
        b = np.array([BASE.index(ch) for ch in s], np.uint32)
        b = np.concatenate([b, np.zeros((-len(b)) % 16, np.uint32)]).reshape(-1, 16)
        words.append((b << (2 * np.arange(16, dtype=np.uint32))).sum(axis=1, dtype=np.uint64).astype(np.uint32))
    return np.concatenate(words), np.array([len(s) for s in seqs], np.uint32)


def pyr(ch):
    return 0 if ch in "CG" else 1


def unit_of(payload, u):
    return "".join(BASE[(payload >> (2 * i)) & 3] for i in range(u))


MH_BASE = {2: 72, 3: 73, 4: 75, 5: 78}
MH_MAX = {2: 1, 3: 2, 4: 3, 5: 5}


def brute_sites(s, zero=True):
    """Sites per ID class of one chromosome as (q, payload), read off the
    bases as a string: every left-normalised one-unit deletion and insertion
    (the unit follows the anchor base s[q - 1]), classified by the copies of
    the unit in the 3' flank and the deleted unit's microhomology with it."""
    out = defaultdict(list)
    n = len(s)
    for q in range(n):
        for u in range(1, 6):
            if q < u or q + u > n:
                continue
            unit, before = s[q:q + u], s[q - u:q]
            if zero:
                if u == 1:
                    for b, pair in ((0, "CG"), (1, "TA")):
                        for bit, ins in enumerate(pair):
                            if ins != unit and ins != before:
                                out[12 + 6 * b].append((q, bit))
                else:
                    for pl in range(4 ** u):
                        ins = unit_of(pl, u)
                        if ins != unit and ins != before:
                            out[48 + 6 * (u - 2)].append((q, pl))
            if unit == before:
                continue
            k = 1
            while s.startswith(unit, q + k * u):
                k += 1
            if u == 1:
                out[6 * pyr(unit) + min(k, 6) - 1].append((q, 0))
                out[12 + 6 * pyr(unit) + min(k, 5)].append((q, 0))
                continue
            if k >= 2:
                out[24 + 6 * (u - 2) + min(k, 6) - 1].append((q, 0))
            else:
                h = min(len(os.path.commonprefix([unit, s[q + u:q + 2 * u]])), u - 1)
                out[MH_BASE[u] + min(h, MH_MAX[u]) - 1 if h else 24 + 6 * (u - 2)].append((q, 0))
            out[48 + 6 * (u - 2) + min(k, 5)].append((q, 0))
    return out


def crafted(rng, n):
    """Random bases with runs of every unit length and copy number spliced in."""
    parts = []
    for u in range(1, 6):
        for k in (1, 2, 3, 5, 6, 8):
            parts.append("".join(rng.choice(list(BASE), n)))
            parts.append("".join(rng.choice(list(BASE), u)) * k)
    return "".join(parts)


def test_id_sites_match_brute_force_small():
    rng = np.random.default_rng(5)
    seqs = [crafted(rng, 9), "GATTACA" + crafted(rng, 4)]
    packed, lens = pack(seqs)
    ix = O.id_index(lens, packed)
    bf = [brute_sites(s) for s in seqs]
    for k in range(NC):
        exp = sum(len(b[k]) for b in bf)
        assert int(ix["counts"][k]) == exp, k
    assert (ix["counts"] > 0).sum() > 50
    # every brute-force site is a site for the oracle; perturbed ones that brute force does not list are not
    for c, b in enumerate(bf):
        for k, sites in b.items():
            step = max(1, len(sites) // 40)
            for q, pl in sites[::step]:
                assert O.id_site(lens, packed, c, q, pl, k), (c, q, pl, k)
        allsites = {(k, q, pl) for k, v in b.items() for q, pl in v}
        for _ in range(3000):
            k = int(rng.integers(NC))
            q = int(rng.integers(len(seqs[c])))
            pl = int(rng.integers(4 ** 5)) if k in (48, 54, 60, 66) else int(rng.integers(2)) if k in (12, 18) else 0
            pl %= max(1, {48: 16, 54: 64, 60: 256, 66: 1024, 12: 2, 18: 2}.get(k, 1))
            assert O.id_site(lens, packed, c, q, pl, k) == ((k, q, pl) in allsites), (c, q, pl, k)


def test_id_site_lists_match_brute_force():
    """A 320 kb reference: classes with few sites (rejection would average
    over 1,024 tries) have their sites listed, in position order."""
    rng = np.random.default_rng(11)
    seqs = ["".join(rng.choice(list(BASE), 200_000)) + crafted(rng, 500),
            "".join(rng.choice(list(BASE), 90_000))]
    packed, lens = pack(seqs)
    ix = O.id_index(lens, packed)
    bf = [brute_sites(s, zero=False) for s in seqs]
    G = int(lens.sum())
    listed = np.nonzero(ix["listed"])[0]
    assert len(listed) >= 10
    boff = [0, int(lens[0])]
    for k in range(NC):
        if k in (12, 18, 48, 54, 60, 66):
            continue  # insertions next to 0 copies: dense, counted in the small test
        assert int(ix["counts"][k]) == sum(len(b[k]) for b in bf), k
        want = ix["counts"][k] > 0 and int(ix["counts"][k]) * 1024 < G
        assert bool(ix["listed"][k]) == bool(want), k
        if ix["listed"][k]:
            exp = sorted(((boff[c] + q) << 10) | pl for c, b in enumerate(bf) for q, pl in b[k])
            got = ix["list"][int(ix["off"][k]):int(ix["off"][k + 1])].tolist()
            assert got == exp, k


def test_id_counts_closed_forms():
    """On an i.i.d. uniform reference a position is a run start of a base of
    either pair with probability 3/8, the run has exactly n bases with
    probability (1/4)^(n-1) 3/4; a 2 bp unit occurs once with 3' microhomology
    of one base with probability 15/16 x 1/4 x 3/4; a 1 bp insertion next to
    0 copies has 2 x (3/4)^2 payloads per position on average."""
    lens = np.array([2_000_000], np.uint32)
    packed = O.genome_synth(lens, 77)
    ix = O.id_index(lens, packed)
    N = float(lens[0])
    cnt = ix["counts"].astype(np.float64)

    def close(k, p):
        mean = N * p
        assert abs(cnt[k] - mean) < 6 * np.sqrt(mean) + 5, (k, cnt[k], mean)

    for b in (0, 1):
        for n in range(1, 6):
            close(6 * b + n - 1, 0.5 * 0.75 * 0.25 ** (n - 1) * 0.75)
        close(6 * b + 5, 0.5 * 0.75 * 0.25 ** 5)
        close(12 + 6 * b, 2 * 0.75 ** 2)
    close(72, 15 / 16 * 0.25 * 0.75)
    # 2 bp units repeated exactly twice: run start, one more copy, then a different unit
    close(24 + 1, 15 / 16 * (1 / 16) * (15 / 16))


def id_signature(rng):
    return rng.dirichlet(np.ones(NC))


def classify_record(seqs, r):
    """ID class of a placed indel, read off the reference as a string: the
    unit right after the anchor base, the copies of it from there on and its
    microhomology with the 3' flank (None when the indel is on no site)."""
    s = seqs[int(r["chr"])]
    q, u = int(r["pos"]) + 1, int(r["len"])
    unit, here, before = unit_of(int(r["seq"]), u), s[q:q + u], s[q - u:q]
    if q < u or q + u > len(s):
        return None
    if r["kind"] == A.PSX_INS and unit != here:
        if unit == before:
            return None
        return 12 + 6 * pyr(unit) if u == 1 else 48 + 6 * (u - 2)  # next to 0 copies
    if unit != here or before == here:
        return None
    k = 1
    while s.startswith(unit, q + k * u):
        k += 1
    if r["kind"] == A.PSX_INS:
        return 12 + 6 * pyr(unit) + min(k, 5) if u == 1 else 48 + 6 * (u - 2) + min(k, 5)
    if u == 1:
        return 6 * pyr(unit) + min(k, 6) - 1
    if k >= 2:
        return 24 + 6 * (u - 2) + min(k, 6) - 1
    h = min(len(os.path.commonprefix([unit, s[q + u:q + 2 * u]])), u - 1)
    return MH_BASE[u] + min(h, MH_MAX[u]) - 1 if h else 24 + 6 * (u - 2)


def test_placement_from_id_signatures():
    """3,000 pre-neoplastic indels from one ID signature: every indel deletes
    (or inserts) its unit right after the anchor at a site of a class, the
    class histogram matches the signature renormalised over the classes with
    a site (chi-square), the cause names the signature after the SBS ones."""
    rng = np.random.default_rng(3)
    seqs = ["".join(rng.choice(list(BASE), 150_000)) + crafted(rng, 300), "".join(rng.choice(list(BASE), 60_000))]
    packed, lens = pack(seqs)
    fd = cases.forest(n_cells=200, n_samples=1, sample_size=40, seed=2)
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    sig = id_signature(rng)
    eng = cases.engine(packed, lens, snv=0.0, indel=0.0, pre_snv=0, pre_indel=3000, drivers=False, germline=False,
                       infinite_sites=False)
    eng.update(id_sig=sig[None, :], id_exposure=np.ones((len(eng["epoch_start"]), 1)), pre_id=0)
    recs = O.place_mutations(lens, packed, of, 17, **eng)
    ind = recs[recs["kind"] != A.PSX_SNV]
    assert len(ind) == 3000
    assert (ind["cause"] == eng["sbs"].shape[0]).all()
    ix = O.id_index(lens, packed)
    hist = np.zeros(NC)
    for r in ind:
        q = int(r["pos"]) + 1
        s = seqs[int(r["chr"])]
        assert s[q - 1] == BASE[r["ref"]]
        if r["kind"] == A.PSX_DEL:
            assert unit_of(int(r["seq"]), int(r["len"])) == s[q:q + int(r["len"])]
        k = classify_record(seqs, r)
        assert k is not None, r
        hist[k] += 1
    p = np.where(ix["counts"] > 0, sig, 0.0)
    p /= p.sum()
    assert (hist[ix["counts"] == 0] == 0).all()
    keep = p * len(ind) >= 5
    exp = p[keep] * len(ind)
    obs = hist[keep]
    chi = ((obs - exp) ** 2 / exp).sum() + ((hist[~keep].sum() - p[~keep].sum() * len(ind)) ** 2 /
                                             max(p[~keep].sum() * len(ind), 1e-9))
    dof = int(keep.sum())
    assert stats.chi2.sf(chi, dof) > 0.001, (chi, dof)


def test_id_signature_errors():
    lens = np.array([5_000], np.uint32)
    packed = O.genome_synth(lens, 1)
    fd = cases.forest(n_cells=60, n_samples=1, sample_size=10, seed=2)
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    eng = cases.engine(packed, lens, snv=0.0, indel=1e-5, pre_snv=0, pre_indel=1, drivers=False, germline=False)
    sig = np.zeros(NC)
    sig[47] = 1.0  # 5 bp units repeated 6+ times: no such site in 5 kb of uniform bases
    eng.update(id_sig=sig[None, :], id_exposure=np.ones((len(eng["epoch_start"]), 1)), pre_id=0)
    with pytest.raises(O.OracleError, match="no class with a site"):
        O.place_mutations(lens, packed, of, 1, **eng)
    eng.update(id_sig=np.ones((1, NC)) / NC, pre_id=3)
    with pytest.raises(O.OracleError, match="unknown preneoplastic ID signature"):
        O.place_mutations(lens, packed, of, 1, **eng)
