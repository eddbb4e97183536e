"""Read materialisation with copy-number events in the oracle: on a
chromosome with events only the accepted templates have reads, in template
order. Pinned against the oracle's count-only path, which draws the same
templates: the alt and depth recounted from the rows (observed bases for
SNVs, the source lineage for indels) equal its counts for every single
(output sample, chromosome) run, the rows number its accepted reads, every
row's source holds a copy of its lineage at the read start (no read from a
deleted lineage), and a chromosome without events keeps the plain rows."""
from __future__ import annotations

import numpy as np
import pytest

import cases
from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core

LENS = np.array([60_000, 25_000], np.uint32)
L = 100
SEQ = dict(coverage=25, read_size=L, insert_size_mean=250, insert_size_stddev=10, error_rate=4e-3, purity=0.7)


@pytest.fixture(scope="module")
def world():
    packed = O.genome_synth(LENS, 21)
    fd = core.synth_forest(growth=[0.2, 0.45], death=[0.01, 0.01], intro_at=[0, 20], n_cells=400, n_samples=2,
                           sample_size=80, seed=5)
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    eng = cases.engine(packed, LENS, snv=6e-6, indel=1e-6, pre_snv=20, pre_indel=3, drivers=False, germline=True)
    eng["cna"] = [(0, 0, 5_000, 20_000, 0, A.PSX_CNA_AMP, 2), (0, 0, 30_000, 15_000, 1, A.PSX_CNA_DEL, 0),
                  (1, 0, 0, 0, 0, A.PSX_CNA_WGD, 0)]
    recs = O.place_mutations(LENS, packed, of, 33, **eng)
    ev = O.place_cna(LENS, of, **eng)
    return dict(packed=packed, of=of, recs=recs, ev=ev)


def profile(ev, leaf, x, ploidy=2):
    cn = [1] * ploidy
    for e in ev:
        if not (e["leaf_lo"] <= leaf < e["leaf_hi"]):
            continue
        if e["kind"] == A.PSX_CNA_WGD:
            cn = [2 * c for c in cn]
        elif e["chr"] == 0 and e["start"] <= x < e["end"]:
            cn[e["allele"]] = cn[e["allele"]] + e["copies"] if e["kind"] == A.PSX_CNA_AMP else 0
    return cn


def recount(recs, rows, chr_, n_leaves, preneo_in_normal=False):
    """alt and depth of every record from materialised reads of one chromosome."""
    M = len(recs)
    alt, dep = np.zeros(M, np.int64), np.zeros(M, np.int64)
    ids = np.nonzero(recs["chr"] == chr_)[0]
    mp = recs["pos"][ids].astype(np.int64)

    def carried(r, leaf, allele):
        if r["allele"] != allele:
            return False
        if leaf == n_leaves:
            return r["nature"] == A.PSX_GERMINAL or (r["nature"] == A.PSX_PRENEOPLASTIC and preneo_in_normal)
        return r["leaf_lo"] <= leaf < r["leaf_hi"]

    for row in rows:
        x, leaf, allele = int(row[0]), int(row[1] & 0x3FFFFFF), int((row[1] >> 26) & 0xF)
        for k in range(np.searchsorted(mp, x), np.searchsorted(mp, x + L)):
            q = ids[k]
            off = int(recs["pos"][q]) - x
            base = int((row[2 + off // 16] >> (2 * (off % 16))) & 3)
            dep[q] += 1
            if not recs["sid_head"][q]:
                continue
            if recs["kind"][q] == A.PSX_SNV:
                alt[q] += base == recs["alt"][q]
            else:
                j, car = q, False
                while j < M and (j == q or not recs["sid_head"][j]):
                    car |= recs["pos"][j] == recs["pos"][q] and carried(recs[j], leaf, allele)
                    j += 1
                alt[q] += car
    return alt, dep


@pytest.mark.parametrize("out", ["sample0", "sample1", "normal"])
@pytest.mark.parametrize("chr_", [0, 1])
def test_rows_reproduce_counts(world, out, chr_):
    of, recs, ev = world["of"], world["recs"], world["ev"]
    seq = dict(SEQ, chromosomes=[chr_])
    if out == "normal":
        seq.update(samples=[], with_normal_sample=True)
    else:
        seq.update(samples=[int(out[-1])], with_normal_sample=False)
    a, d, _, n = O.simulate_seq(LENS, of, recs, 17, cna=ev, **seq)
    rows = O.simulate_reads(LENS, world["packed"], of, recs, 17, 2 * n + 10, cna=ev, **seq)
    assert len(rows) == n
    ra, rd = recount(recs, rows, chr_, of.n_leaves)
    assert np.array_equal(rd, d[0]) and np.array_equal(ra, a[0])
    if chr_ == 0 and out != "normal":
        # every read comes from a copy its source holds at the read start
        sources = {(int(r[1] & 0x3FFFFFF), int((r[1] >> 26) & 0xF), int(r[0])) for r in rows[::2]}
        for leaf, allele, x in sources:
            if leaf < of.n_leaves:
                assert profile(ev, leaf, x)[allele] > 0


def test_chromosome_without_events_keeps_plain_rows(world):
    """Events on chromosome 0 only: chromosome 1's rows equal the plain rows
    (no events at all) of a run over chromosome 1, and chromosome 0's rows
    number its accepted reads."""
    of, recs, ev = world["of"], world["recs"], world["ev"]
    assert set(ev["chr"].tolist()) <= {0, A.PSX_ALL_CHROMOSOMES}
    ev0 = ev[ev["kind"] != A.PSX_CNA_WGD]  # a doubling touches every chromosome
    both = O.simulate_reads(LENS, world["packed"], of, recs, 9, 10**6, cna=ev0, **SEQ)
    _, _, _, n0 = O.simulate_seq(LENS, of, recs, 9, cna=ev0,
                                 **dict(SEQ, chromosomes=[0], samples=[0], with_normal_sample=False))
    plain1 = O.simulate_reads(LENS, world["packed"], of, recs, 9, 10**6,
                              **dict(SEQ, chromosomes=[1], samples=[0], with_normal_sample=False))
    # output sample 0's segments come first: chromosome 0 (with events), then chromosome 1
    assert np.array_equal(both[n0:n0 + len(plain1)], plain1)


def test_capacity_counts_accepted_reads(world):
    of, recs, ev = world["of"], world["recs"], world["ev"]
    _, _, _, n = O.simulate_seq(LENS, of, recs, 5, cna=ev, **SEQ)
    with pytest.raises(O.OracleError, match="the read buffer holds"):
        O.simulate_reads(LENS, world["packed"], of, recs, 5, n - 1, cna=ev, **SEQ)
    assert len(O.simulate_reads(LENS, world["packed"], of, recs, 5, n, cna=ev, **SEQ)) == n
