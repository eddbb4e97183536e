"""Checkpoint and resume: a phylogenetic forest saved and loaded back
(PhylogeneticForest.save / load_phylogenetic_forest over the library's
mutation-table format, psx_mutations_save / psx_mutations_load). On the CPU:
the loader's refusals that come before any device work, with the reference
loader's messages (src/phylogenetic_forest.cpp:958-986,
src/utility.cpp:141-161). On the GPU: round trips (records, copy-number
events, the samples forest, and simulate_seq's whole table from the loaded
forest) and the table file's own checks (magic, size, checksum, version,
records out of range)."""
from __future__ import annotations

import json
import os
import struct
import zipfile

import numpy as np
import pytest

import cases
import process_amd as pa
from process_amd import _abi as A
from process_amd import core
from process_amd._lib import PsxError


# ---------------------------------------------------------------- CPU refusals

def test_forest_file_refusals(tmp_path):
    missing = tmp_path / "none.pf"
    with pytest.raises(ValueError, match=f'^The file "{missing}" does not exist.$'):
        pa.load_phylogenetic_forest(missing)
    with pytest.raises(ValueError, match="is not a regular file"):
        pa.load_phylogenetic_forest(tmp_path)
    text = tmp_path / "text.pf"
    text.write_text("not a forest")
    with pytest.raises(ValueError, match=f'^Wrong file format for the phylogenetic forest file "{text}". Remove it'):
        pa.load_phylogenetic_forest(text)
    other = tmp_path / "other.pf"
    with zipfile.ZipFile(other, "w") as z:
        z.writestr("meta.json", json.dumps({"format": "something else", "version": 1}))
    with pytest.raises(ValueError, match="Wrong file format"):
        pa.load_phylogenetic_forest(other)
    newer = tmp_path / "newer.pf"
    with zipfile.ZipFile(newer, "w") as z:
        z.writestr("meta.json", json.dumps({"format": "psx phylogenetic forest", "version": 7}))
    with pytest.raises(ValueError, match=f'"{newer}" is in the file format version 7, expected version 1. Remove it'):
        pa.load_phylogenetic_forest(newer)
    partial = tmp_path / "partial.pf"
    with zipfile.ZipFile(partial, "w") as z:
        z.writestr("meta.json", json.dumps({"format": "psx phylogenetic forest", "version": 1, "seed": 1}))
    with pytest.raises(ValueError, match="Wrong file format"):
        pa.load_phylogenetic_forest(partial)


# ------------------------------------------------------------------ GPU trips

def build(seed=3):
    sig = cases.signatures(2, 11)
    ref = pa.ReferenceGenome({"21": 60_000, "22": 45_000}, seed=17)
    eng = pa.MutationEngine(ref, {"SBS1": sig[0], "SBS5": sig[1]}, germline=[("22", 1001, "T", 1)],
                            context_sampling=1)
    eng.add_mutant("A", {"SNV": 5e-6, "indel": 5e-7})
    eng.add_mutant("B", {"SNV": 8e-6, "indel": 5e-7},
                   drivers=[pa.SNV("21", 30_001, "A"), pa.CNA("D", "22", 2_001, 9_000, allele=0), pa.WGD])
    eng.add_exposure({"SBS1": 0.7, "SBS5": 0.3})
    sf = pa.SamplesForest.synthetic(mutants={"A": (0.2, 0.01), "B": (0.3, 0.01)}, introductions={"B": 100},
                                    n_cells=500, samples={"S1": 100, "S2": 100}, seed=seed)
    return eng, sf


@pytest.mark.gpu
def test_forest_round_trip(ctx, tmp_path):
    eng, sf = build()
    pf = eng.place_mutations(sf, 10, 3, seed=42)
    path = tmp_path / "forest.pf"
    pf.save(path)
    back = pa.load_phylogenetic_forest(path)
    assert back.records().tobytes() == pf.records().tobytes()
    assert back.mutations.cna().tobytes() == pf.mutations.cna().tobytes() and len(pf.mutations.cna()) >= 2
    assert back.seed == pf.seed and back.mutations.ploidy == 2
    assert back.engine.reference.names == eng.reference.names
    assert back.engine.signature_names == eng.signature_names
    assert back.get_samples_info().equals(pf.get_samples_info())
    kw = dict(sequencer=pa.BasicIlluminaSequencer(1e-3), coverage=30, insert_size_mean=350, insert_size_stddev=10,
              seed=7)
    a, b = pa.simulate_seq(pf, **kw)["mutations"], pa.simulate_seq(back, **kw)["mutations"]
    assert a.equals(b) and len(a) > 20


def fnv_words(data: bytes) -> int:
    data = data + b"\0" * (-len(data) % 8)
    h = 0xcbf29ce484222325
    for (w,) in struct.iter_unpack("<Q", data):
        h = ((h ^ w) * 0x100000001b3) & 0xFFFFFFFFFFFFFFFF
    return h


@pytest.mark.gpu
def test_table_file_checks(ctx, tmp_path):
    lens = [20_000, 9_000]
    g = core.Genome(ctx, lens, seed=2)
    fd = cases.forest(n_cells=300, n_samples=1, sample_size=60, seed=4)
    f = core.Forest(ctx, fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])
    eng = cases.engine(g.download(), lens, snv=2e-5, indel=2e-6)
    m = core.Mutations.place(ctx, g, f, 9, **eng)
    path = tmp_path / "t.psx"
    m.save(path)
    back = core.Mutations.load(ctx, path)
    assert back.download().tobytes() == m.download().tobytes() and back.n == m.n > 50
    raw = path.read_bytes()

    def refused(data: bytes, match: str):
        bad = tmp_path / "bad.psx"
        bad.write_bytes(data)
        with pytest.raises(PsxError, match=match):
            core.Mutations.load(ctx, bad)

    fmt = "Wrong file format for the mutation table file"
    refused(b"XSXMUTBL" + raw[8:], fmt)                                   # magic
    refused(raw[:-1], fmt)                                                # size
    flipped = bytearray(raw)
    flipped[200] ^= 1
    refused(bytes(flipped), fmt)                                          # checksum
    refused(raw[:8] + struct.pack("<I", 2) + raw[12:], "in the file format version 2, expected version 1")
    # a record out of range under a valid checksum: allele 5 of a diploid table
    body = bytearray(raw[:-8])
    n_chr = struct.unpack_from("<I", body, 20)[0]
    first = 40 + 4 * n_chr
    rec = np.frombuffer(bytes(body[first:first + A.MUTATION_DTYPE.itemsize]), A.MUTATION_DTYPE).copy()
    rec["allele"] = 5
    body[first:first + A.MUTATION_DTYPE.itemsize] = rec.tobytes()
    refused(bytes(body) + struct.pack("<Q", fnv_words(bytes(body))), fmt)
    # and the unmodified body with its recomputed checksum loads (the checksum is the one written)
    assert fnv_words(raw[:-8]) == struct.unpack("<Q", raw[-8:])[0]
    with pytest.raises(PsxError, match="does not exist"):
        core.Mutations.load(ctx, tmp_path / "missing.psx")
    os.mkdir(tmp_path / "dir.psx")
    with pytest.raises(PsxError, match="is not a regular file"):
        core.Mutations.load(ctx, tmp_path / "dir.psx")
