"""The committed sequencing fixtures (tests/golden/seq_*.npz, written by
tests/golden/make_golden.py from a brute-force read-by-read model): the
oracle reproduces them on the CPU, the gfx950 path on the GPU, both through
the mutation-table replay entry point (the fixture's records are the input)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from process_amd import _abi as A

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem[len("seq_"):] for p in GOLDEN.glob("seq_*.npz"))


def load(case):
    z = np.load(GOLDEN / f"seq_{case}.npz")
    d = {k: z[k] for k in z.files}
    d["recs"] = d["recs"].view(A.MUTATION_DTYPE)
    d["cna"] = d["cna"].view(A.CNA_EVENT_DTYPE) if "cna" in d else np.zeros(0, A.CNA_EVENT_DTYPE)
    d["seq"] = json.loads(str(d["seq"]))
    return d


def test_fixtures_present():
    assert set(CASES) >= {"pe_finite_sites", "dense_finite_sites", "se_subset", "cna_purity"}


@pytest.mark.parametrize("case", CASES)
def test_fixture_shapes(case):
    d = load(case)
    S, M = d["alt"].shape
    assert M == len(d["recs"]) and d["depth"].shape == (S, M) and d["present"].shape == (S, M)
    # alt reads never exceed depth, and only SID heads collect them
    assert (d["alt"] <= d["depth"]).all()
    assert (d["alt"][:, d["recs"]["sid_head"] == 0] == 0).all()


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_fixture(case):
    d = load(case)
    of = O.Forest(d["cell_id"], d["parent"], d["species"], d["sample"], d["birth"], int(d["n_samples"]))
    a, dp, pr, n = O.simulate_seq(d["lens"], of, d["recs"], int(d["seed"]), cna=d["cna"], **d["seq"])
    assert n == int(d["n_reads"])
    assert np.array_equal(a, d["alt"]) and np.array_equal(dp, d["depth"]) and np.array_equal(pr, d["present"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_matches_fixture(ctx, case):
    from process_amd import core

    print(f"[golden] {case}", flush=True)
    d = load(case)
    g = core.Genome(ctx, d["lens"], packed=d["packed"])
    f = core.Forest(ctx, d["cell_id"], d["parent"], d["species"], d["sample"], d["birth"], int(d["n_samples"]))
    m = core.Mutations.upload(ctx, f, d["recs"])
    m.set_cna(d["cna"])
    assert np.array_equal(m.download(), d["recs"])
    a, dp, pr, n = core.simulate_seq_arrays(ctx, g, f, m, int(d["seed"]), **d["seq"])
    assert n == int(d["n_reads"])
    assert np.array_equal(a, d["alt"]) and np.array_equal(dp, d["depth"]) and np.array_equal(pr, d["present"])


READ_CASES = [c for c in CASES if c != "cna_purity"]  # reads are materialised without copy-number events


@pytest.mark.parametrize("case", READ_CASES)
def test_oracle_reads_match_fixture(case):
    """Every read, bit for bit: start, source leaf, allele, mate, error flag
    and 2-bit bases (reference, carried SNVs, sequencing errors)."""
    d = load(case)
    assert len(d["reads"]) == int(d["n_reads"]) > 0
    of = O.Forest(d["cell_id"], d["parent"], d["species"], d["sample"], d["birth"], int(d["n_samples"]))
    rows = O.simulate_reads(d["lens"], d["packed"], of, d["recs"], int(d["seed"]), len(d["reads"]), **d["seq"])
    assert np.array_equal(rows, d["reads"])
    # errors flagged on some reads only, and reads of both mates when paired
    flag = d["reads"][:, 1] >> 31
    assert 0 < flag.sum() < len(flag)
    if d["seq"].get("insert_size_mean", 0) > 0:
        assert ((d["reads"][:, 1] >> 30) & 1).tolist() == [0, 1] * (len(d["reads"]) // 2)


@pytest.mark.gpu
@pytest.mark.parametrize("case", READ_CASES)
def test_gpu_reads_match_fixture(ctx, case):
    from process_amd import core

    print(f"[golden] reads {case}", flush=True)
    d = load(case)
    g = core.Genome(ctx, d["lens"], packed=d["packed"])
    f = core.Forest(ctx, d["cell_id"], d["parent"], d["species"], d["sample"], d["birth"], int(d["n_samples"]))
    m = core.Mutations.upload(ctx, f, d["recs"])
    out, n = core.simulate_seq_reads(ctx, g, f, m, int(d["seed"]), **d["seq"])
    assert n == len(d["reads"])
    assert np.array_equal(out.numpy()[:n], d["reads"])


def hand_table():
    """Four records: SID A held by two records (head + member), SID B, an
    indel SID C; two samples. Expected rows follow get_active_mutations and
    add_sample_statistics (src/seq_simulation.cpp:92-167): occurrences count
    only where the sample holds the SID, VAF = occ/cov there and 0 elsewhere."""
    recs = np.zeros(4, A.MUTATION_DTYPE)
    recs["pos"] = [100, 100, 250, 400]
    recs["sid_head"] = [1, 0, 1, 1]
    recs["kind"] = [A.PSX_SNV, A.PSX_SNV, A.PSX_SNV, A.PSX_DEL]
    alt = np.array([[3, 0, 0, 0], [0, 0, 0, 2]], np.int32)
    dep = np.array([[10, 10, 7, 5], [8, 8, 0, 4]], np.int32)
    pres = np.array([[1, 0, 0, 0], [0, 1, 1, 0]], np.uint8)
    return recs, alt, dep, pres


@pytest.mark.parametrize("impl", ["oracle", "libpsx"])
def test_result_table_hand_fixture(impl):
    from process_amd import core

    fn = O.seq_table if impl == "oracle" else core.seq_table
    recs, alt, dep, pres = hand_table()
    rows, occ, cov, vaf = fn(recs, alt, dep, pres, False)
    assert rows.tolist() == [0, 3]
    assert occ.tolist() == [[3, 0], [0, 2]]
    assert cov.tolist() == [[10, 5], [8, 4]]
    assert vaf.tolist() == [[0.3, 0.0], [0.0, 0.5]]
    rows, occ, cov, vaf = fn(recs, alt, dep, pres, True)
    assert rows.tolist() == [0, 2, 3]
    assert occ.tolist() == [[3, 0, 0], [0, 0, 2]]
    assert cov.tolist() == [[10, 7, 5], [8, 0, 4]]
    assert vaf[0].tolist() == [0.3, 0.0, 0.0]
    assert vaf[1, 0] == 0.0 and np.isnan(vaf[1, 1]) and vaf[1, 2] == 0.5


@pytest.mark.parametrize("case", READ_CASES)
@pytest.mark.parametrize("threads", [1, 3])
def test_reference_shaped_restatement_matches(case, threads):
    """The reference-shaped CPU restatement (read strings, ordered maps) gives
    the fixture's counts: the CPU baseline it times does the same work."""
    d = load(case)
    of = O.Forest(d["cell_id"], d["parent"], d["species"], d["sample"], d["birth"], int(d["n_samples"]))
    a, dp, n = O.refshape_seq(d["lens"], d["packed"], of, d["recs"], int(d["seed"]), threads, **d["seq"])
    assert n == int(d["n_reads"])
    assert np.array_equal(a, d["alt"]) and np.array_equal(dp, d["depth"])
