"""The R-style API: argument validation and table reshaping on the CPU; the
end-to-end path (MutationEngine -> place_mutations -> simulate_seq) on the GPU
against the oracle and the result-table semantics of the reference."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

import cases
import process_amd as pa
from oracle import oracle as O
from process_amd import _abi as A


def test_sequencer_validation_messages():
    """src/sequencers.cpp:86-101."""
    with pytest.raises(ValueError, match="must be a positive real number"):
        pa.BasicIlluminaSequencer(-1e-3)
    with pytest.raises(ValueError, match="must be a positive real number"):
        pa.BasicIlluminaSequencer("x")
    with pytest.raises(ValueError, match="must be a Boolean value"):
        pa.BasicIlluminaSequencer(1e-3, random_quality_scores=1)
    s = pa.BasicIlluminaSequencer(4e-3)
    assert s.error_rate == 4e-3 and s.random_quality_scores
    assert pa.ErrorlessIlluminaSequencer().error_rate == 0.0


def test_seq_to_long_reshapes_like_the_reference():
    """R/seq_to_long.R:31-67 on the example of its roxygen block."""
    wide = pd.DataFrame({"chr": ["chr1", "chr2"], "chr_pos": [100, 200], "ref": ["A", "C"], "alt": ["T", "G"],
                         "causes": ["SBS5", "SBS1"], "classes": ["germinal", "passenger"],
                         "Sample.A.occurrences": [10, 90], "Sample.A.coverage": [100, 100],
                         "Sample.A.VAF": [0.1, 0.9], "normal_sample.occurrences": [45, 52],
                         "normal_sample.coverage": [100, 100], "normal_sample.VAF": [0.45, 0.52]})
    long = pa.seq_to_long({"mutations": wide})
    assert len(long) == 4
    assert list(long["sample_name"]) == ["Sample.A", "Sample.A", "normal_sample", "normal_sample"]
    assert list(long["NV"]) == [10, 90, 45, 52] and list(long["DP"]) == [100] * 4
    assert list(long["from"]) == list(long["to"]) == [100, 200, 100, 200]


def _dr_baf_table():
    return pd.DataFrame({"chr": ["1", "2", "1", "2"], "chr_pos": [100, 50, 30, 10], "ref": ["A", "C", "G", "T"],
                         "alt": ["T", "G", "A", "C"], "causes": ["SBS1"] * 4, "classes": ["germinal"] * 4,
                         "S.occurrences": [30, 0, 5, 1], "S.coverage": [60, 0, 20, 4],
                         "S.VAF": [0.5, 0.0, 0.25, 0.25], "normal_sample.occurrences": [15, 0, 4, 2],
                         "normal_sample.coverage": [30, 0, 0, 8], "normal_sample.VAF": [0.5, 0.0, 0.0, 0.25]})


def test_depth_ratio_data_like_plot_dr():
    """R/plot_genome_wide_mutations.R:75-111 without the display subsample:
    (chr, from) order by chromosome level, DR = DP.tumour / DP.normal."""
    d = pa.seq_dr({"mutations": _dr_baf_table()}, "S")
    assert list(zip(d["chr"], d["from"])) == [("1", 30), ("1", 100), ("2", 10), ("2", 50)]
    assert list(d["abs_pos"]) == [1, 2, 3, 4]
    assert d["DR"][0] == np.inf and d["DR"][1] == 2.0 and d["DR"][2] == 0.5 and np.isnan(d["DR"][3])
    assert list(d["NV.tumour"]) == [5, 30, 1, 0] and list(d["NV.normal"]) == [4, 15, 2, 0]
    # chromosome order given by the caller, one chromosome left out
    d2 = pa.seq_dr(_dr_baf_table(), "S", chromosomes=["2"])
    assert list(d2["from"]) == [10, 50] and list(d2["DR"][:1]) == [0.5]
    d3 = pa.seq_dr(_dr_baf_table(), "S", chromosomes=["2", "1"])
    assert list(zip(d3["chr"], d3["from"])) == [("2", 10), ("2", 50), ("1", 30), ("1", 100)]


def test_baf_data_like_plot_baf():
    """R/plot_genome_wide_mutations.R:192-240 without the display subsample:
    abs_pos is numbered before the VAF cut."""
    b = pa.seq_baf(_dr_baf_table(), "S", cuts=(0.4, 0.1))
    assert list(b["from"]) == [30, 10] and list(b["abs_pos"]) == [1, 3]
    assert len(pa.seq_baf(_dr_baf_table(), "S")) == 4


def test_dr_baf_errors_like_the_reference():
    """R/ggplot_config.R:82-90, R/plot_genome_wide_mutations.R:12-15,87-94."""
    t = _dr_baf_table()
    with pytest.raises(ValueError, match="^The chromosome X is not present in the sequence reference data.$"):
        pa.seq_baf(t, "S", chromosomes=["1", "X"])
    with pytest.raises(ValueError, match="^The chromosomes X, Y are not present"):
        pa.seq_dr(t, "S", chromosomes=["X", "Y", "X"])
    with pytest.raises(ValueError, match="available samples are: S, normal_sample$"):
        pa.seq_baf(t, "T")
    with pytest.raises(ValueError, match="mandatory normal sample"):
        pa.seq_dr(t.drop(columns=["normal_sample.VAF"]), "S")


def test_random_seed():
    assert pa.get_random_seed(5) == 5
    s = pa.get_random_seed(None)
    assert -(2**31) <= s < 2**31
    with pytest.raises(TypeError):
        pa.get_random_seed(1.5)


def test_engine_rejects_unknown_signature_and_rates():
    ref = pa.ReferenceGenome({"1": 1000})
    with pytest.raises(ValueError, match="96 classes"):
        pa.MutationEngine(ref, {"SBS1": [1.0] * 10})
    eng = pa.MutationEngine(ref, {"SBS1": cases.signatures(1, 0)[0]})
    # src/mutation_engine.cpp:917-928
    with pytest.raises(ValueError, match='^"SNP" is an unsupported name in passenger rates list.$'):
        eng.add_mutant("A", {"SNP": 1e-8})
    with pytest.raises(ValueError, match="CNA rates must be non-negative"):
        eng.add_mutant("A", {"SNV": 1e-8, "CNA": -1.0})
    # src/cna.cpp:100-138
    with pytest.raises(ValueError, match='Unknown CNA type "X"'):
        pa.CNA("X", "1", 10, 100)
    with pytest.raises(ValueError, match="Region lengths must be non-negative numbers"):
        pa.Deletion("1", 10, -5)
    eng.add_mutant("B", {"SNV": 1e-8, "CNA": 1e-10}, drivers=[pa.WGD, pa.Amplification("1", 11, 40, src_allele=0)])
    sf = pa.SamplesForest([1, 2], [-1, 1], ["B", "B"], [None, "S"], [0.0, 1.0])
    eng.add_exposure({"SBS1": 1.0})
    with pytest.raises(ValueError, match="need a passenger CNA catalogue"):
        eng._params(sf, 0, "SBS1", 0)
    with pytest.raises(ValueError, match="no known SBS signature"):
        eng.add_exposure({"SBS5": 1.0})
    with pytest.raises(ValueError, match="sum to 1"):
        eng.add_exposure({"SBS1": 0.5})


def test_read_buffer_shape_is_checked():
    """A caller's read buffer narrower than a read, or of another dtype, is
    refused before the library writes rows into it."""
    from types import SimpleNamespace

    from process_amd import core
    core._check_read_buffer(SimpleNamespace(dtype=np.dtype(np.uint32), shape=(10, 12)), 12)
    for shape, dt in [((10, 11), np.uint32), ((10, 12), np.int64), ((120,), np.uint32)]:
        with pytest.raises(ValueError, match="read buffer"):
            core._check_read_buffer(SimpleNamespace(dtype=np.dtype(dt), shape=shape), 12)


# ------------------------------------------------------------------- GPU

def build(seed=3):
    sig = cases.signatures(2, 11)
    ref = pa.ReferenceGenome({"21": 60_000, "22": 45_000}, seed=17)
    eng = pa.MutationEngine(ref, {"SBS1": sig[0], "SBS5": sig[1]}, germline=[("22", 1001, "T", 1)],
                            context_sampling=1)
    eng.add_mutant("A", {"SNV": 5e-6, "indel": 5e-7})
    eng.add_mutant("B", {"SNV": 8e-6, "indel": 5e-7}, drivers=[pa.SNV("21", 30_001, "A")])
    eng.add_exposure({"SBS1": 0.7, "SBS5": 0.3})
    eng.add_exposure(6.0, {"SBS5": 1.0})
    sf = pa.SamplesForest.synthetic(mutants={"A": (0.2, 0.01), "B": (0.3, 0.01)}, introductions={"B": 100},
                                    n_cells=500, samples={"S1": 100, "S2": 100}, seed=seed)
    return eng, sf


@pytest.mark.gpu
def test_end_to_end_against_oracle(ctx):
    eng, sf = build()
    if eng.base(0, 30_000) == 0:
        pytest.skip("driver alt equals the reference base")
    pf = eng.place_mutations(sf, 10, 3, seed=42)
    recs = pf.records()
    # the same placement on the oracle
    packed = eng._packed
    of = O.Forest(sf.cell_id, sf.ancestor, sf.species_index(), sf.sample_index(sf.sample_names), sf.birth_time, 2)
    om = O.place_mutations(eng.reference.lengths, packed, of, 42,
                           **eng._params(sf, 10, "SBS1", 3))
    assert recs.tobytes() == om.tobytes()
    res = pa.simulate_seq(pf, sequencer=pa.BasicIlluminaSequencer(1e-3), coverage=40, insert_size_mean=350,
                          insert_size_stddev=10, seed=7)
    df = res["mutations"]
    assert res["parameters"]["seed"] == 7
    oa, od, op, on = O.simulate_seq(eng.reference.lengths, of, om, 7, coverage=40, read_size=150,
                                    insert_size_mean=350, insert_size_stddev=10, error_rate=1e-3)
    rows, occ, cov, vaf = O.seq_table(om, oa, od, op)
    assert len(df) == len(rows)
    for s, name in enumerate(["S1", "S2", "normal_sample"]):
        assert np.array_equal(df[f"{name}.occurrences"].to_numpy(), occ[s])
        assert np.array_equal(df[f"{name}.coverage"].to_numpy(), cov[s])
        assert np.allclose(df[f"{name}.VAF"].to_numpy(), vaf[s], equal_nan=True)
    assert (df["chr_pos"].to_numpy() == om["pos"][rows.astype(np.int64)] + 1).all()
    # the germline SNV is in the normal sample at about half the reads
    g = df[(df["chr"] == "22") & (df["chr_pos"] == 1001)]
    assert len(g) == 1 and g["classes"].iloc[0] == "germinal"
    assert 0.2 < g["normal_sample.VAF"].iloc[0] < 0.8
    assert set(df["classes"]) <= {"passenger", "driver", "germinal", "preneoplastic"}
    # the DR / BAF consumers on the simulated table: one row per table row, depths taken from it
    dr = pa.seq_dr(res, "S1")
    assert len(dr) == len(df) and list(dr["abs_pos"]) == list(range(1, len(df) + 1))
    key = dict(zip(zip(df["chr"], df["chr_pos"]), zip(df["S1.coverage"], df["normal_sample.coverage"])))
    assert all(key[(c, f)] == (t, n) for c, f, t, n in zip(dr["chr"], dr["from"], dr["DP.tumour"], dr["DP.normal"]))
    assert len(pa.seq_baf(res, "S2", chromosomes=["22"])) == int((df["chr"] == "22").sum())


@pytest.mark.gpu
def test_normal_seq_labelling_and_sam(ctx, tmp_path):
    eng, sf = build(seed=5)
    pf = eng.place_mutations(sf, 5, 0, seed=1)
    res = pa.simulate_normal_seq(pf, coverage=20, write_SAM=False, seed=3)
    cols = [c for c in res["mutations"].columns if c.endswith(".VAF")]
    assert cols == ["normal_sample.VAF"]
    assert set(res["mutations"]["classes"]) <= {"germinal", "preneoplastic"}
    # FACS labelling splits the samples by label
    lab = pa.simulate_seq(pf, coverage=10, cell_labelling=lambda c: "early" if c.birth_time < 20 else "late",
                          with_normal_sample=False, include_non_sequenced_mutations=True, seed=2)
    names = sorted(c[: -4] for c in lab["mutations"].columns if c.endswith(".VAF"))
    assert all(n.startswith(("S1_", "S2_")) for n in names)
    # SAM files: one per chromosome, reads of every sample as read groups
    out = tmp_path / "sam"
    pa.simulate_seq(pf, sequencer=pa.BasicIlluminaSequencer(1e-3), coverage=2, insert_size_mean=350,
                    write_SAM=True, output_dir=str(out), seed=4)
    files = sorted(p.name for p in out.iterdir())
    assert files == ["chr_21.sam", "chr_22.sam"]
    lines = (out / "chr_22.sam").read_text().splitlines()
    rg = [l for l in lines if l.startswith("@RG")]
    assert len(rg) == 3
    body = [l.split("\t") for l in lines if not l.startswith("@")]
    assert len(body) == 3 * 2 * round(2 * 45_000 / 300)
    assert all(len(f[9]) == 150 and f[5] == "150M" for f in body)
    ref = eng.reference
    # mismatches against the reference are rare (SNVs of the source cell and 1e-3 errors)
    packed = eng._packed
    off = (60_000 + 15) // 16
    mism = 0
    for f in body[:200]:
        x = int(f[3]) - 1
        for i, ch in enumerate(f[9]):
            p = x + i
            b = (packed[off + p // 16] >> (2 * (p % 16))) & 3
            mism += "ACGT"[b] != ch
    assert mism < 0.01 * 200 * 150
    assert ref.names == ["21", "22"]
