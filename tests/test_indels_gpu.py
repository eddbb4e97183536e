"""Indel placement from ID83 signatures on the GPU against the oracle: the
genome's indel index (classes counted per site, sparse classes listed) built
on the device and used by the placement kernel gives the oracle's table byte
for byte, with dense classes (rejection) and listed classes both drawn, under
the infinite-sites model and with timed exposures; sequencing over those
indels matches too; and the engine API attributes indels to ID signatures."""
from __future__ import annotations

import numpy as np
import pytest

import cases
import process_amd as pa
from oracle import oracle as O
from process_amd import _abi as A
from process_amd import core

pytestmark = pytest.mark.gpu

BASE = "ACGT"


def crafted(rng, n):
    parts = []
    for u in range(1, 6):
        for k in (1, 2, 3, 5, 6, 8):
            parts.append("".join(rng.choice(list(BASE), n)))
            parts.append("".join(rng.choice(list(BASE), u)) * k)
    return "".join(parts)


def genome(rng):
    seqs = ["".join(rng.choice(list(BASE), 180_000)) + crafted(rng, 800),
            crafted(rng, 300) + "".join(rng.choice(list(BASE), 70_000))]
    ref = pa.ReferenceGenome({"1": len(seqs[0]), "2": len(seqs[1])}, sequences={"1": seqs[0], "2": seqs[1]})
    return ref, seqs


def signatures(ix, rng):
    """Two ID signatures: one Dirichlet over all classes, one concentrated on
    listed (sparse) classes that have sites."""
    s0 = rng.dirichlet(np.ones(A.PSX_ID_CLASSES))
    s1 = np.zeros(A.PSX_ID_CLASSES)
    listed = np.nonzero(ix["listed"] & (ix["counts"] > 0))[0]
    s1[listed] = 1.0
    s1[0] = s1[12] = 0.5
    return np.array([s0, s1 / s1.sum()]), listed


def test_id_placement_matches_oracle(ctx):
    rng = np.random.default_rng(8)
    ref, seqs = genome(rng)
    lens = ref.lengths
    packed = ref.packed
    g = core.Genome(ctx, lens, packed=packed)
    fd = cases.forest(n_cells=900, n_samples=2, sample_size=150, seed=5)
    f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    of = O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    ix = O.id_index(lens, packed)
    sigs, listed = signatures(ix, rng)
    assert len(listed) >= 5
    eng = cases.engine(packed, lens, snv=2e-6, indel=2e-6, pre_snv=4, pre_indel=40)
    eng.update(id_sig=sigs, id_exposure=[[0.8, 0.2], [0.1, 0.9]], pre_id=1)
    m = core.Mutations.place(ctx, g, f, 1234, **eng)
    om = O.place_mutations(lens, packed, of, 1234, **eng)
    got = m.download()
    assert got.tobytes() == om.tobytes()
    ind = om[om["kind"] != A.PSX_SNV]
    n_sbs = eng["sbs"].shape[0]
    assert len(ind) > 200 and set(ind["cause"].tolist()) == {n_sbs, n_sbs + 1}
    # listed classes were drawn: some indels sit inside the crafted repeats (unit lengths 2..5, 5+ copies)
    assert (ind["len"] >= 2).sum() > 20
    seq = dict(coverage=30, read_size=150, insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3)
    a, d, p, n = core.simulate_seq_arrays(ctx, g, f, m, 6, **seq)
    oa, od, op, on = O.simulate_seq(lens, of, om, 6, **seq)
    assert n == on and np.array_equal(a, np.reshape(oa, a.shape)) and np.array_equal(d, np.reshape(od, d.shape))
    assert np.array_equal(p, np.reshape(op, p.shape))


def test_id_signatures_through_the_engine(ctx):
    """MutationEngine with ID signatures: SBS and ID exposures in one list,
    a later time changing the SBS part only, pre-neoplastic indels from a
    named ID signature; the placement equals the oracle's and the result
    table names the ID signatures in its causes."""
    rng = np.random.default_rng(21)
    ref, seqs = genome(rng)
    ix = O.id_index(ref.lengths, ref.packed)
    sigs, _ = signatures(ix, rng)
    sbs = cases.signatures(2, 4)
    eng = pa.MutationEngine(ref, {"SBS1": sbs[0], "SBS5": sbs[1]}, ID_signatures={"ID1": sigs[0], "ID8": sigs[1]},
                            context_sampling=1)
    eng.add_mutant("A", {"SNV": 2e-6, "indel": 1e-6})
    eng.add_mutant("B", {"SNV": 3e-6, "indel": 2e-6})
    eng.add_exposure({"SBS1": 0.6, "SBS5": 0.4, "ID1": 0.7, "ID8": 0.3})
    eng.add_exposure(8.0, {"SBS5": 1.0})
    sf = pa.SamplesForest.synthetic(mutants={"A": (0.2, 0.01), "B": (0.3, 0.01)}, introductions={"B": 80},
                                    n_cells=700, samples={"S1": 100}, seed=6)
    pf = eng.place_mutations(sf, 3, 20, seed=55, preneoplastic_indel_signature_name="ID8")
    params = eng._params(sf, 3, "SBS1", 20, "ID8")
    assert params["pre_id"] == 1
    assert np.allclose(params["id_exposure"], [[0.7, 0.3], [0.7, 0.3]])
    assert np.allclose(params["exposure"], [[0.6, 0.4], [0.0, 1.0]])
    of = O.Forest(sf.cell_id, sf.ancestor, sf.species_index(), sf.sample_index(sf.sample_names), sf.birth_time, 1)
    om = O.place_mutations(ref.lengths, ref.packed, of, 55, **params)
    assert pf.records().tobytes() == om.tobytes()
    res = pa.simulate_seq(pf, coverage=20, include_non_sequenced_mutations=True, seed=3)
    causes = set(res["mutations"]["causes"].dropna())
    assert {"ID1", "ID8"} <= causes and {"SBS1", "SBS5"} & causes
    with pytest.raises(ValueError, match='Unknown ID signature "ID9"'):
        eng.place_mutations(sf, 0, 2, seed=1, preneoplastic_indel_signature_name="ID9")
