"""R-level API of ProCESS's mutation engine and sequencing simulation, running
on the gfx950 path (libpsx).

Names, arguments, defaults and result layout follow the reference's Rcpp
modules: ``MutationEngine$place_mutations`` (src/mutations.cpp:546-573,
src/mutation_engine.hpp:160-206), ``simulate_seq`` / ``simulate_normal_seq``
(src/sequencing.cpp:193-209, 268-282; src/seq_simulation.cpp:517-679),
``BasicIlluminaSequencer`` / ``ErrorlessIlluminaSequencer``
(src/sequencing.cpp:36-128, src/sequencers.cpp:80-106) and ``seq_to_long``
(R/seq_to_long.R:31-67). Inputs that the reference downloads (reference
FASTA, COSMIC signatures) are given explicitly or synthesised.
"""
from __future__ import annotations

import os
import random
from dataclasses import dataclass
from typing import Callable, Iterable, Mapping, Sequence

import numpy as np
import pandas as pd

from . import _abi as A
from . import core

BASES = np.array(list("ACGT"))
NATURE = {A.PSX_PASSENGER: "passenger", A.PSX_DRIVER: "driver", A.PSX_PRENEOPLASTIC: "preneoplastic",
          A.PSX_GERMINAL: "germinal"}


def get_random_seed(seed) -> int:
    """src/utility.hpp:41-64: an integer seed, or one drawn when NULL."""
    if seed is None:
        return random.randrange(-(2**31), 2**31)
    if isinstance(seed, (bool, np.bool_)) or not isinstance(seed, (int, np.integer)):
        raise TypeError("The seed must be an integer or None.")
    return int(seed)


def _seed64(seed: int) -> int:
    return int(seed) & 0xFFFFFFFFFFFFFFFF


# ---------------------------------------------------------------- inputs

class SamplesForest:
    """Genealogy of the sampled cells (SamplesForest, src/samples_forest.hpp:31-82).

    One row per cell: ``cell_id``, ``ancestor`` (-1 for roots), ``mutant``
    (species name), ``sample`` (name of a sampled leaf, None otherwise),
    ``birth_time``."""

    def __init__(self, cell_id, ancestor, mutant: Sequence[str], sample: Sequence, birth_time,
                 epistate: Sequence[str] | None = None):
        self.cell_id = np.asarray(cell_id, np.int64)
        self.ancestor = np.asarray(ancestor, np.int64)
        self.mutant = np.asarray(mutant, dtype=object)
        self.epistate = np.asarray(epistate if epistate is not None else [""] * len(self.cell_id), dtype=object)
        self.sample = np.asarray(sample, dtype=object)
        self.birth_time = np.asarray(birth_time, np.float64)
        names = [s for s in dict.fromkeys(self.sample.tolist()) if s is not None and s == s]
        self.sample_names = names
        species = list(dict.fromkeys(zip(self.mutant.tolist(), self.epistate.tolist())))
        self.species = species

    @classmethod
    def synthetic(cls, *, mutants: Mapping[str, tuple[float, float]], introductions: Mapping[str, int] | None = None,
                  n_cells: int, samples: Mapping[str, int], seed: int) -> "SamplesForest":
        """A seeded birth-death genealogy standing in for
        ``SpatialSimulation$get_samples_forest()`` (the spatial simulation stays
        outside the accelerated path). ``mutants`` maps a name to (growth,
        death); the first founds the tumour; ``introductions`` gives, for the
        others, the population size at which one founder cell converts."""
        names = list(mutants)
        intro = introductions or {}
        sizes = list(samples.values())
        if len(set(sizes)) != 1:
            raise ValueError("synthetic samples must have equal sizes")
        fd = core.synth_forest(growth=[mutants[n][0] for n in names], death=[mutants[n][1] for n in names],
                               intro_at=[0] + [int(intro.get(n, 2**62)) for n in names[1:]], n_cells=n_cells,
                               n_samples=len(samples), sample_size=sizes[0], seed=seed)
        snames = list(samples)
        sample = [None if s < 0 else snames[s] for s in fd["sample"].tolist()]
        return cls(fd["cell_id"], fd["parent"], [names[s] for s in fd["species"]], sample, fd["birth"])

    def get_nodes(self) -> pd.DataFrame:
        """src/forest.hpp:29-67."""
        return pd.DataFrame({"cell_id": self.cell_id, "ancestor": np.where(self.ancestor < 0, pd.NA, self.ancestor),
                             "mutant": self.mutant, "epistate": self.epistate, "sample": self.sample,
                             "birth_time": self.birth_time})

    def get_samples_info(self) -> pd.DataFrame:
        counts = pd.Series(self.sample).value_counts()
        return pd.DataFrame({"name": self.sample_names,
                             "tumour_cells": [int(counts.get(n, 0)) for n in self.sample_names]})

    def species_index(self) -> np.ndarray:
        idx = {s: i for i, s in enumerate(self.species)}
        return np.array([idx[(m, e)] for m, e in zip(self.mutant.tolist(), self.epistate.tolist())], np.uint32)

    def sample_index(self, names: Sequence[str]) -> np.ndarray:
        idx = {n: i for i, n in enumerate(names)}
        return np.array([idx.get(s, -1) if s is not None else -1 for s in self.sample.tolist()], np.int32)


@dataclass
class SNV:
    """A driver SNV specification (src/sid.hpp:28-77); chr_pos is 1-based."""
    chr: str
    chr_pos: int
    alt: str
    ref: str | None = None
    allele: int | None = None


class CNA:
    """A copy-number alteration driver, CNA(type, chr, chr_pos, len, allele,
    src_allele) (src/mutations.cpp:198-262, src/cna.cpp:100-138): type "A"
    (amplification of the src_allele lineage) or "D" (deletion of the allele
    lineage); chr_pos is 1-based; an allele left None is drawn uniformly at
    placement."""

    def __init__(self, type: str, chr, chr_pos, len, allele=None, src_allele=None):
        if type not in ("A", "D"):
            raise ValueError(f"Unknown CNA type \"{type}\". Supported types are \"A\" and \"D\" for amplification "
                             "and deletion, respectively")
        if int(chr_pos) < 0:
            raise ValueError("Position in chromosome must be a non-negative number")
        if int(len) < 0:
            raise ValueError("Region lengths must be non-negative numbers")
        self.type, self.chr, self.chr_pos, self.len = type, str(chr), int(chr_pos), int(len)
        self.allele = None if allele is None else int(allele)
        self.src_allele = None if src_allele is None else int(src_allele)

    def lineage(self):
        """The allele lineage the event acts on: the source of an amplification, the allele of a deletion."""
        return self.src_allele if self.type == "A" else self.allele

    def __repr__(self):
        return f"CNA(\"{self.type}\", \"{self.chr}\", {self.chr_pos}, {self.len})"


def Amplification(chr, chr_pos, len, allele=None, src_allele=None) -> CNA:
    """src/mutations.cpp:224-240."""
    return CNA("A", chr, chr_pos, len, allele, src_allele)


def Deletion(chr, chr_pos, len, allele=None) -> CNA:
    """src/mutations.cpp:244-258."""
    return CNA("D", chr, chr_pos, len, allele)


class _WGD:
    """Whole-genome doubling driver (src/wg_doubling.hpp, src/mutations.cpp:355-365)."""

    def __repr__(self):
        return "WGD"


WGD = _WGD()


def load_passenger_CNAs(path: str, tumour_type: str = "", tumour_study: str = "") -> list[tuple[str, int, int, str]]:
    """The passenger CNA catalogue of a tab-separated file with a header
    (load_passenger_CNAs, src/mutation_engine.cpp:397-487): columns chr,
    begin, end, major, minor, tumour type, study; rows of other tumour types
    or studies are skipped when those are given; a row gives an amplification
    when major is NA or > 1 and a deletion when minor is NA or < 1, over the
    1-based region [begin, end]. Returns (chr, begin, end, "A"/"D") tuples,
    without duplicates, in (chr, begin, length, type) order."""
    out = set()
    with open(path) as f:
        lines = f.read().splitlines()
    for row_num, line in enumerate(lines[1:], start=2):
        if not line.strip():
            continue
        row = line.split("\t")
        if len(row) < 7:
            raise ValueError("The CNA CSV must contains at least 7 columns")
        if (tumour_type and row[5] != tumour_type) or (tumour_study and row[6] != tumour_study):
            continue
        for k, what in ((1, "begin"), (2, "end")):
            if not row[k].strip().isdigit():
                raise ValueError(f"Unknown {what} specification {row[k]} in row number {row_num}.")
        begin, end = int(row[1]), int(row[2])
        if begin > end:
            raise ValueError(f"The CNA begin lays after the end in row number {row_num}")
        major, minor = row[3], row[4]
        try:
            if major == "NA" or int(major) > 1:
                out.add((row[0], begin, end, "A"))
        except ValueError:
            raise ValueError(f"Unknown major specification {major} in row number {row_num}.") from None
        try:
            if minor == "NA" or int(minor) < 1:
                out.add((row[0], begin, end, "D"))
        except ValueError:
            raise ValueError(f"Unknown minor specification {major} in row number {row_num}.") from None
    return sorted(out, key=lambda r: (r[0], r[1], r[2] - r[1], r[3]))


class ReferenceGenome:
    """Chromosome names and lengths with a 2-bit reference: synthetic i.i.d.
    uniform ACGT from a seed, or bases given explicitly."""

    def __init__(self, chromosomes: Mapping[str, int], seed: int = 0, sequences: Mapping[str, str] | None = None):
        self.names = list(chromosomes)
        self.lengths = np.array([int(chromosomes[n]) for n in self.names], np.uint32)
        self.seed = int(seed)
        self.packed = None
        if sequences is not None:
            words = []
            code = {c: i for i, c in enumerate("ACGT")}
            for n in self.names:
                s = sequences[n].upper()
                b = np.array([code.get(c, 0) for c in s], np.uint32)
                b = np.concatenate([b, np.zeros((-len(b)) % 16, np.uint32)]).reshape(-1, 16)
                words.append((b << (2 * np.arange(16, dtype=np.uint32))).sum(axis=1, dtype=np.uint64).astype(np.uint32))
            self.packed = np.concatenate(words)

    def chr_index(self, name: str) -> int:
        try:
            return self.names.index(str(name))
        except ValueError:
            raise ValueError(f"Unknown chromosome \"{name}\"") from None


class MutationEngine:
    """The mutation engine (src/mutation_engine.hpp:35-52): species passenger
    rates (SNV, indel, CNA), drivers (SNVs, CNAs, whole-genome doublings),
    timed SBS and ID exposures, the passenger CNA catalogue and the
    infinite-sites flag."""

    def __init__(self, reference: ReferenceGenome, SBS_signatures: Mapping[str, Sequence[float]],
                 germline: Iterable[tuple[str, int, str, int]] = (), ctx: core.Context | None = None,
                 indel_max_len: int = 10, indel_len_p: float = 0.5,
                 ID_signatures: Mapping[str, Sequence[float]] | None = None, passenger_CNAs=None,
                 tumour_type: str = "", tumour_study: str = "", context_sampling: int = 100):
        """ID_signatures: COSMIC ID83 signatures (83 classes each); without
        them indels have geometric lengths and uniform positions.
        context_sampling: the SBS context index keeps every
        context_sampling-th site of each trinucleotide context
        (src/mutations.cpp:818-819, default 100; 1 keeps every site).
        passenger_CNAs: the passenger CNA catalogue, a tab-separated file
        (load_passenger_CNAs, filtered by tumour_type and tumour_study) or
        (chr, begin, end, "A"/"D") tuples with 1-based inclusive regions."""
        self._ctx = ctx
        self.reference = reference
        self.signature_names = list(SBS_signatures)
        self.signatures = np.array([np.asarray(SBS_signatures[n], np.float64) for n in self.signature_names])
        if self.signatures.shape[1:] != (96,):
            raise ValueError("SBS signatures must have 96 classes")
        self.id_signature_names = list(ID_signatures or {})
        self.id_signatures = np.array([np.asarray(ID_signatures[n], np.float64) for n in self.id_signature_names]) \
            if ID_signatures else np.zeros((0, A.PSX_ID_CLASSES))
        if self.id_signatures.shape[1:] != (A.PSX_ID_CLASSES,):
            raise ValueError(f"ID signatures must have {A.PSX_ID_CLASSES} classes")
        overlap = set(self.signature_names) & set(self.id_signature_names)
        if overlap:
            raise ValueError(f"signature names used twice: {sorted(overlap)}")
        if isinstance(passenger_CNAs, (str, os.PathLike)):
            passenger_CNAs = load_passenger_CNAs(str(passenger_CNAs), tumour_type, tumour_study)
        self.passenger_CNAs = [(str(c), int(b), int(e), str(t)) for c, b, e, t in (passenger_CNAs or [])]
        self.mutants: dict[str, dict] = {}
        self.exposures: list[tuple[float, dict]] = []
        self.infinite_sites_model = True
        self.germline = list(germline)
        self.indel_max_len = indel_max_len
        self.indel_len_p = indel_len_p
        if int(context_sampling) != context_sampling or context_sampling < 1:
            raise ValueError("context_sampling must be a positive integer")
        self.context_sampling = int(context_sampling)
        self._genome: core.Genome | None = None

    @property
    def ctx(self) -> core.Context:
        if self._ctx is None:
            self._ctx = core.default_context()
        return self._ctx

    def genome(self) -> core.Genome:
        if self._genome is None:
            r = self.reference
            self._genome = core.Genome(self.ctx, r.lengths, seed=r.seed, packed=r.packed)
            self._packed = self._genome.download()
        return self._genome

    def base(self, chr_idx: int, pos0: int) -> int:
        self.genome()
        off = int(sum((int(l) + 15) // 16 for l in self.reference.lengths[:chr_idx]))
        return int((self._packed[off + pos0 // 16] >> (2 * (pos0 % 16))) & 3)

    def add_mutant(self, mutant_name: str, passenger_rates: Mapping, drivers: Sequence = ()):
        """src/mutation_engine.cpp:1055-1083 and :901-944: ``passenger_rates``
        is c(SNV=, indel=, CNA=) or a per-epistate list of them; ``drivers``
        holds SNVs, CNAs and WGD, applied in the given order."""
        rates = dict(passenger_rates)
        if rates and all(isinstance(v, Mapping) for v in rates.values()):
            per_epi = {k: dict(v) for k, v in rates.items()}
        else:
            per_epi = {"": rates}
        for r in per_epi.values():
            for k, v in r.items():
                if k not in ("SNV", "indel", "CNA"):
                    raise ValueError(f"\"{k}\" is an unsupported name in passenger rates list.")
                if not (float(v) >= 0.0):
                    raise ValueError(f": {k} rates must be non-negative")
        for d in drivers:
            if not isinstance(d, (SNV, CNA, _WGD)):
                raise ValueError(f"unsupported driver {d!r}: SNV, CNA or WGD")
        self.mutants[mutant_name] = {"rates": per_epi, "drivers": list(drivers)}

    def add_exposure(self, time_or_coefficients, coefficients: Mapping[str, float] | None = None):
        """src/mutation_engine.cpp:734-739: exposure from a simulated time on;
        SBS and ID coefficients may share one list, each part summing to 1. A
        part left out keeps the exposure of the earlier time."""
        if coefficients is None:
            time, coefficients = 0.0, time_or_coefficients
        else:
            time = float(time_or_coefficients)
        sbs = {k: float(v) for k, v in coefficients.items() if k in self.signature_names}
        ids = {k: float(v) for k, v in coefficients.items() if k in self.id_signature_names}
        unknown = [k for k in coefficients if k not in sbs and k not in ids]
        if unknown or not (sbs or ids):
            raise ValueError("The exposure names no known SBS signature" if not (sbs or ids) else
                             f"Unknown signature {unknown[0]}")
        for part in (sbs, ids):
            if part and abs(sum(part.values()) - 1.0) > 1e-6:
                raise ValueError("The SBS exposure coefficients must sum to 1" if part is sbs else
                                 "The ID exposure coefficients must sum to 1")
        coef = dict(sbs)
        coef.update(ids)
        self.exposures = [e for e in self.exposures if e[0] != time] + [(time, coef)]
        self.exposures.sort(key=lambda e: e[0])

    def get_species_rates(self) -> pd.DataFrame:
        rows = []
        for m, d in self.mutants.items():
            for epi, r in d["rates"].items():
                rows.append({"mutant": m, "epistate": epi, "SNV": r.get("SNV", 0.0), "indel": r.get("indel", 0.0),
                             "CNA": r.get("CNA", 0.0)})
        return pd.DataFrame(rows)

    def _epoch_exposures(self, names: list[str]) -> tuple[list[float], np.ndarray] | None:
        """Per exposure time, the coefficients over `names`: a time whose list
        names none of them keeps the previous time's (None if no time does)."""
        times, rows, cur = [], [], None
        for t, coef in self.exposures:
            part = [coef.get(n, 0.0) for n in names] if any(n in coef for n in names) else None
            if part is not None:
                cur = part
            times.append(t)
            rows.append(cur)
        if cur is None:
            return None
        first = next(r for r in rows if r is not None)
        return times, np.array([r if r is not None else first for r in rows], np.float64)

    def _params(self, forest: SamplesForest, n_pre_snv: int, pre_sig: str, n_pre_indel: int,
                pre_id_sig: str = "ID1") -> dict:
        snv, ind, cna = [], [], []
        for mutant, epi in forest.species:
            if mutant not in self.mutants:
                raise ValueError(f"The mutant \"{mutant}\" has no passenger rates in the mutation engine")
            rates = self.mutants[mutant]["rates"]
            r = rates.get(epi, rates.get("", {}))
            snv.append(float(r.get("SNV", 0.0)))
            ind.append(float(r.get("indel", 0.0)))
            cna.append(float(r.get("CNA", 0.0)))
        if n_pre_snv and pre_sig not in self.signature_names:
            raise ValueError(f"Unknown SBS signature \"{pre_sig}\"")
        sbs_expo = self._epoch_exposures(self.signature_names)
        if sbs_expo is None and (any(snv) or n_pre_snv):
            raise ValueError("No SBS exposure has been added to the mutation engine")
        use_id = len(self.id_signature_names) > 0
        id_expo = self._epoch_exposures(self.id_signature_names) if use_id else None
        if use_id:
            if n_pre_indel and pre_id_sig not in self.id_signature_names:
                raise ValueError(f"Unknown ID signature \"{pre_id_sig}\"")
            if id_expo is None and any(ind):
                raise ValueError("No ID exposure has been added to the mutation engine")
        if any(cna) and not self.passenger_CNAs:
            raise ValueError("Passenger CNA rates need a passenger CNA catalogue")
        sp_index = {s: i for i, s in enumerate(forest.species)}
        drivers, cna_drivers = [], []
        for mutant, d in self.mutants.items():
            species = [s for (m, e), s in sp_index.items() if m == mutant]
            for spec in d["drivers"]:
                if isinstance(spec, SNV):
                    c = self.reference.chr_index(spec.chr)
                    pos0 = int(spec.chr_pos) - 1
                    if spec.ref is not None and BASES[self.base(c, pos0)] != spec.ref:
                        raise ValueError(f"The reference base of {spec} does not match the reference genome")
                    for s in species:
                        drivers.append((s, c, pos0, "ACGT".index(spec.alt),
                                        -1 if spec.allele is None else int(spec.allele)))
                elif isinstance(spec, CNA):
                    c = self.reference.chr_index(spec.chr)
                    lin = spec.lineage()
                    for s in species:
                        cna_drivers.append((s, c, max(spec.chr_pos, 1) - 1, spec.len,
                                            A.PSX_CNA_RANDOM_ALLELE if lin is None else lin,
                                            A.PSX_CNA_AMP if spec.type == "A" else A.PSX_CNA_DEL,
                                            1 if spec.type == "A" else 0))
                else:
                    for s in species:
                        cna_drivers.append((s, 0, 0, 0, 0, A.PSX_CNA_WGD, 0))
        germ = [(self.reference.chr_index(c), int(p) - 1, "ACGT".index(a), int(al)) for c, p, a, al in self.germline]
        # one epoch per exposure time: SBS and ID parts each from the latest time that set them
        epochs = [t for t, _ in self.exposures] or [0.0]
        n_ep = len(epochs)
        expo = sbs_expo[1] if sbs_expo is not None else np.ones((n_ep, len(self.signature_names))) / \
            max(1, len(self.signature_names))
        out = dict(snv_rate=snv, indel_rate=ind, sbs=self.signatures, epoch_start=epochs, exposure=expo,
                   n_pre_snv=n_pre_snv, n_pre_indel=n_pre_indel,
                   pre_sbs=self.signature_names.index(pre_sig) if pre_sig in self.signature_names else 0,
                   indel_max_len=self.indel_max_len, indel_len_p=self.indel_len_p,
                   infinite_sites=bool(self.infinite_sites_model), ploidy=2, drivers=drivers, germline=germ,
                   cna=cna_drivers, context_sampling=self.context_sampling)
        if any(cna):
            cat = [(self.reference.chr_index(c), b - 1, e + 1 - b, A.PSX_CNA_AMP if t == "A" else A.PSX_CNA_DEL)
                   for c, b, e, t in self.passenger_CNAs]
            out.update(cna_rate=cna, cna_catalogue=cat)
        if use_id:
            idx = id_expo[1] if id_expo is not None else np.ones((n_ep, len(self.id_signature_names))) / \
                len(self.id_signature_names)
            out.update(id_sig=self.id_signatures, id_exposure=idx,
                       pre_id=self.id_signature_names.index(pre_id_sig) if pre_id_sig in self.id_signature_names
                       else 0)
        return out

    def place_mutations(self, samples_forest: SamplesForest, num_of_preneoplastic_SNVs: int = 0,
                        num_of_preneoplastic_indels: int = 0, seed=None,
                        preneoplastic_SNV_signature_name: str = "SBS1",
                        preneoplastic_indel_signature_name: str = "ID1") -> "PhylogeneticForest":
        """src/mutations.cpp:546-573 and src/mutation_engine.cpp:1085-1113."""
        c_seed = get_random_seed(seed)
        params = self._params(samples_forest, int(num_of_preneoplastic_SNVs), preneoplastic_SNV_signature_name,
                              int(num_of_preneoplastic_indels), preneoplastic_indel_signature_name)
        g = self.genome()
        f = core.Forest(self.ctx, samples_forest.cell_id, samples_forest.ancestor, samples_forest.species_index(),
                        samples_forest.sample_index(samples_forest.sample_names), samples_forest.birth_time,
                        len(samples_forest.sample_names))
        m = core.Mutations.place(self.ctx, g, f, _seed64(c_seed), **params)
        return PhylogeneticForest(self, samples_forest, f, m, seed=c_seed)


class PhylogeneticForest:
    """A samples forest labelled with mutations (src/phylogenetic_forest.hpp:34-142):
    the flattened forest and the mutation table stay device resident."""

    def __init__(self, engine: MutationEngine, samples_forest: SamplesForest, forest: core.Forest,
                 mutations: core.Mutations, seed: int):
        self.engine = engine
        self.samples_forest = samples_forest
        self.forest = forest
        self.mutations = mutations
        self.seed = seed
        self._recs = None
        self._flat = None

    def records(self) -> np.ndarray:
        if self._recs is None:
            self._recs = self.mutations.download()
        return self._recs

    def flat(self) -> dict:
        if self._flat is None:
            self._flat = self.forest.flat()
        return self._flat

    def mutation_strings(self, recs: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        ref, alt = [], []
        for r in recs:
            b = BASES[r["ref"]]
            seq = "".join(BASES[(int(r["seq"]) >> (2 * i)) & 3] for i in range(int(r["len"])))
            if r["kind"] == A.PSX_SNV:
                ref.append(b); alt.append(BASES[r["alt"]])
            elif r["kind"] == A.PSX_DEL:
                ref.append(b + seq); alt.append(b)
            else:
                ref.append(b); alt.append(b + seq)
        return np.array(ref, dtype=object), np.array(alt, dtype=object)

    def leaf_cells(self) -> np.ndarray:
        """Cell ids of the forest's leaves in leaf-rank (DFS) order."""
        fl = self.flat()
        parent = fl["parent"]
        is_leaf = np.ones(len(parent), bool)
        is_leaf[parent[parent != A.PSX_NO_NODE]] = False
        return fl["cell_id"][np.nonzero(is_leaf)[0]]

    def get_sampled_cell_mutations(self, cell_ids: Iterable[int] | int | None = None) -> pd.DataFrame:
        """Per-cell mutation table (src/phylogenetic_forest.cpp:452-507): a leaf
        carries record m iff leaf_lo(m) <= rank(leaf) < leaf_hi(m). Asking for
        a cell that is not a leaf raises the reference's error (:476-484)."""
        recs = self.records()
        leaf_cells = self.leaf_cells()
        want = None
        if cell_ids is not None:
            want = {int(cell_ids)} if isinstance(cell_ids, (int, np.integer)) else set(int(c) for c in cell_ids)
            known = set(leaf_cells.tolist())
            for c in sorted(want - known):
                raise ValueError(f"Cell \"{c}\" is not a leaf")
        ref, alt = self.mutation_strings(recs)
        chr_names = np.array(self.engine.reference.names, dtype=object)
        rows = []
        for rank, cell in enumerate(leaf_cells.tolist()):
            if want is not None and cell not in want:
                continue
            sel = np.nonzero((recs["leaf_lo"] <= rank) & (rank < recs["leaf_hi"]))[0]
            rows.append(pd.DataFrame({
                "cell_id": cell, "chr": chr_names[recs["chr"][sel]], "chr_pos": recs["pos"][sel].astype(np.int64) + 1,
                "allele": recs["allele"][sel], "ref": ref[sel], "alt": alt[sel],
                "type": np.where(recs["kind"][sel] == A.PSX_SNV, "SNV", "indel"),
                "cause": [self._cause(c) for c in recs["cause"][sel]],
                "class": [NATURE[int(x)] for x in recs["nature"][sel]]}))
        if not rows:
            return pd.DataFrame(columns=["cell_id", "chr", "chr_pos", "allele", "ref", "alt", "type", "cause", "class"])
        return pd.concat(rows, ignore_index=True)

    def _cause(self, c) -> str | None:
        """SBS signature of an SNV; ID signature (index past the SBS ones) of an indel."""
        c = int(c)
        if c == A.PSX_NO_CAUSE:
            return None
        n_sbs = len(self.engine.signature_names)
        return self.engine.signature_names[c] if c < n_sbs else self.engine.id_signature_names[c - n_sbs]

    def get_first_occurrences(self, chr: str, chr_pos: int, ref: str, alt: str) -> list:
        """Cells where the mutation first arises, ascending
        (src/phylogenetic_forest.cpp:564-600): the cell of each record's edge
        node; pre-neoplastic mutations arise in the roots. A mutation held only
        by the germline gives [None] (R's NA); one absent from the sampled
        cells raises the reference's error."""
        recs = self.records()
        c = self.engine.reference.chr_index(chr)
        r_ref, r_alt = self.mutation_strings(recs)
        sel = np.nonzero((recs["chr"] == c) & (recs["pos"] == int(chr_pos) - 1) & (r_ref == ref) & (r_alt == alt))[0]
        fl = self.flat()
        roots = fl["cell_id"][fl["parent"] == A.PSX_NO_NODE]
        cells: set[int] = set()
        germline = False
        for k in sel.tolist():
            nat = int(recs["nature"][k])
            if nat == A.PSX_GERMINAL:
                germline = True
            elif recs["node"][k] != A.PSX_NO_NODE:
                cells.add(int(fl["cell_id"][recs["node"][k]]))
            else:
                cells.update(int(r) for r in roots)
        if cells:
            return sorted(cells)
        if germline:
            return [None]
        raise ValueError(f"The mutation {chr}:{int(chr_pos)} {ref}>{alt} does not occurs in the sampled cells.")

    def get_samples_info(self) -> pd.DataFrame:
        return self.samples_forest.get_samples_info()

    def save(self, filename, quiet: bool = False) -> None:
        """PhylogeneticForest$save (src/mutations.cpp:1446-1460,
        src/phylogenetic_forest.cpp:940-956): the forest in one file, a zip
        archive holding the mutation table in the library's binary format
        (psx_mutations_save), the samples forest, the reference (names,
        lengths, seed or explicit bases), the signatures and the seed; read
        back with load_phylogenetic_forest. quiet is accepted for the
        reference's signature (there is no progress bar)."""
        import io
        import json
        import tempfile
        import zipfile

        sf, ref, eng = self.samples_forest, self.engine.reference, self.engine

        def npz(**arrays) -> bytes:
            b = io.BytesIO()
            np.savez(b, **arrays)
            return b.getvalue()

        text = lambda a: np.array(["" if v is None or v != v else str(v) for v in a], dtype=np.str_)
        with tempfile.TemporaryDirectory() as d:
            table = os.path.join(d, "mutations.psx")
            self.mutations.save(table)
            with zipfile.ZipFile(filename, "w", zipfile.ZIP_STORED) as z:
                z.writestr("meta.json", json.dumps(dict(format=_FOREST_FORMAT, version=_FOREST_VERSION,
                                                        seed=int(self.seed), ploidy=int(self.mutations.ploidy))))
                z.writestr("forest.npz", npz(cell_id=sf.cell_id, ancestor=sf.ancestor, birth_time=sf.birth_time,
                                             mutant=text(sf.mutant), epistate=text(sf.epistate),
                                             sample=text(sf.sample), has_sample=np.array(
                                                 [v is not None and v == v for v in sf.sample], bool)))
                extra = {} if ref.packed is None else dict(packed=ref.packed)
                z.writestr("reference.npz", npz(names=np.array(ref.names, dtype=np.str_), lengths=ref.lengths,
                                                seed=np.array([ref.seed], np.int64), **extra))
                z.writestr("signatures.npz", npz(sbs_names=np.array(eng.signature_names, dtype=np.str_),
                                                 sbs=eng.signatures,
                                                 id_names=np.array(eng.id_signature_names, dtype=np.str_),
                                                 id=eng.id_signatures))
                z.write(table, "mutations.psx")


_FOREST_FORMAT = "psx phylogenetic forest"
_FOREST_VERSION = 1


def load_phylogenetic_forest(filename, quiet: bool = False, ctx: core.Context | None = None) -> PhylogeneticForest:
    """load_phylogenetic_forest (src/mutations.cpp:1466-1478,
    src/phylogenetic_forest.cpp:958-992): a forest written by
    PhylogeneticForest.save, its mutation table back in device memory. The
    engine it carries holds the reference and the signatures (what sequencing
    and the result table use), as the reference's loaded forest holds its
    reference path and exposures. Errors follow the reference's loader."""
    import io
    import json
    import tempfile
    import zipfile

    path = str(filename)
    if not os.path.exists(path):
        raise ValueError(f"The file \"{path}\" does not exist.")
    if not os.path.isfile(path):
        raise ValueError(f"The file \"{path}\" is not a regular file.")
    wrong = ValueError(f"Wrong file format for the phylogenetic forest file \"{path}\". Remove it and retry.")
    try:
        z = zipfile.ZipFile(path, "r")
    except zipfile.BadZipFile:
        raise wrong from None
    with z, tempfile.TemporaryDirectory() as d:
        try:
            meta = json.loads(z.read("meta.json"))
            if meta.get("format") != _FOREST_FORMAT:
                raise wrong
            if meta.get("version") != _FOREST_VERSION:
                raise ValueError(f"The phylogenetic forest file \"{path}\" is in the file format version "
                                 f"{meta.get('version')}, expected version {_FOREST_VERSION}. Remove it and retry.")
            load = lambda name: np.load(io.BytesIO(z.read(name)), allow_pickle=False)
            fz, rz, sz = load("forest.npz"), load("reference.npz"), load("signatures.npz")
            z.extract("mutations.psx", d)
        except (KeyError, ValueError, json.JSONDecodeError) as e:
            if isinstance(e, ValueError) and str(e).startswith(("The phylogenetic", "Wrong file format")):
                raise
            raise wrong from None
        ref = ReferenceGenome(dict(zip(rz["names"].tolist(), rz["lengths"].tolist())), seed=int(rz["seed"][0]))
        if "packed" in rz.files:
            ref.packed = rz["packed"]
        sample = [s if h else None for s, h in zip(fz["sample"].tolist(), fz["has_sample"].tolist())]
        sf = SamplesForest(fz["cell_id"], fz["ancestor"], fz["mutant"].tolist(), sample, fz["birth_time"],
                           epistate=fz["epistate"].tolist())
        eng = MutationEngine(ref, dict(zip(sz["sbs_names"].tolist(), sz["sbs"])), ctx=ctx,
                             ID_signatures=dict(zip(sz["id_names"].tolist(), sz["id"])) or None)
        f = core.Forest(eng.ctx, sf.cell_id, sf.ancestor, sf.species_index(), sf.sample_index(sf.sample_names),
                        sf.birth_time, len(sf.sample_names))
        m = core.Mutations.load(eng.ctx, os.path.join(d, "mutations.psx"))
    return PhylogeneticForest(eng, sf, f, m, seed=int(meta["seed"]))


# --------------------------------------------------------------- sequencers

class ErrorlessIlluminaSequencer:
    """src/sequencers.hpp:25-38."""
    error_rate = 0.0

    def show(self) -> str:
        return "Error-less Illumina sequencer"


class BasicIlluminaSequencer:
    """src/sequencers.hpp:40-72; validation src/sequencers.cpp:80-106."""

    def __init__(self, error_rate, random_quality_scores: bool = True):
        if isinstance(error_rate, bool) or not isinstance(error_rate, (int, float, np.integer, np.floating)):
            raise ValueError("The parameter \"error_rate\" must be a positive real number.")
        if error_rate < 0:
            raise ValueError("The parameter \"error_rate\" must be a positive real number.")
        if not isinstance(random_quality_scores, (bool, np.bool_)):
            raise ValueError("The parameter \"random_quality_scores\" must be a Boolean value.")
        self.error_rate = float(error_rate)
        self.random_quality_scores = bool(random_quality_scores)

    def show(self) -> str:
        kind = "random quality scores" if self.random_quality_scores else "constant quality scores"
        return f"Basic Illumina sequencer (error rate: {self.error_rate} {kind})"


def _sequencer_data(sequencer) -> dict | None:
    if sequencer is None:
        return None
    if isinstance(sequencer, BasicIlluminaSequencer):
        return {"name": "BasicIlluminaSequencer", "error_rate": sequencer.error_rate,
                "random_quality_scores": sequencer.random_quality_scores}
    if isinstance(sequencer, ErrorlessIlluminaSequencer):
        return {"name": "ErrorlessIlluminaSequencer", "error_rate": 0.0}
    raise ValueError("Unsupported sequencer type")


# ---------------------------------------------------------------- sequencing

@dataclass
class SampledCell:
    """The labelling function's argument (src/sampled_cell.hpp)."""
    cell_id: int
    mutant: str
    epistate: str
    species: str
    birth_time: float


def _relevant_chromosomes(reference: ReferenceGenome, chromosomes) -> list[int]:
    """get_relevant_chr_set (src/seq_simulation.cpp:299-352)."""
    if chromosomes is None:
        return list(range(len(reference.names)))
    if isinstance(chromosomes, str):
        chromosomes = [chromosomes]
    out = set()
    for i, c in enumerate(chromosomes):
        if not isinstance(c, str):
            raise ValueError(f"Expected a list of string: the element {i + 1} of the list is not a string.")
        out.add(reference.chr_index(c))
    return sorted(out)


def _labelled_forest(pf: PhylogeneticForest, labelling: Callable) -> tuple[core.Forest, list[str]]:
    """apply_FACS_labels (src/seq_simulation.cpp:183-243): each sample splits
    by the label of its cells into "<sample>_<label>" ("" keeps the name)."""
    sf = pf.samples_forest
    new_sample = np.full(len(sf.cell_id), -1, np.int32)
    names: list[str] = []
    index: dict[str, int] = {}
    for s in sf.sample_names:
        for i in np.nonzero(sf.sample == s)[0]:
            cell = SampledCell(int(sf.cell_id[i]), str(sf.mutant[i]), str(sf.epistate[i]),
                               f"{sf.mutant[i]}{sf.epistate[i]}", float(sf.birth_time[i]))
            label = str(labelling(cell))
            name = s if label == "" else f"{s}_{label}"
            if name not in index:
                index[name] = len(names)
                names.append(name)
            new_sample[i] = index[name]
    f = core.Forest(pf.engine.ctx, sf.cell_id, sf.ancestor, sf.species_index(), new_sample, sf.birth_time, len(names))
    return f, names


def _table(pf: PhylogeneticForest, sample_names: list[str], alt, depth, present,
           include_non_sequenced_mutations: bool) -> pd.DataFrame:
    """get_result_dataframe (src/seq_simulation.cpp:52-181) from psx_seq_table rows."""
    recs = pf.records()
    rows, occ, cov, vaf = core.seq_table(recs, alt, depth, present, include_non_sequenced_mutations)
    rows = rows.astype(np.int64)
    sel = recs[rows]
    ref, alt_s = pf.mutation_strings(sel)
    causes, classes = [], []
    heads = np.nonzero(recs["sid_head"])[0]
    ends = np.append(heads[1:], len(recs))
    end_of = dict(zip(heads.tolist(), ends.tolist()))
    for h in rows.tolist():
        grp = recs[h: end_of[h]]
        cs = sorted({pf._cause(c) for c in grp["cause"]} - {None})
        causes.append(";".join(cs) if cs else None)
        classes.append(";".join(sorted({NATURE[int(x)] for x in grp["nature"]})))
    df = pd.DataFrame({"chr": np.array(pf.engine.reference.names, dtype=object)[sel["chr"]],
                       "chr_pos": sel["pos"].astype(np.int64) + 1, "ref": ref, "alt": alt_s,
                       "causes": causes, "classes": classes})
    for s, name in enumerate(sample_names):
        df[f"{name}.occurrences"] = occ[s]
        df[f"{name}.coverage"] = cov[s]
        df[f"{name}.VAF"] = vaf[s]
    return df


def _run(pf: PhylogeneticForest, forest: core.Forest, sample_names: list[str], sequencer, chromosomes, coverage,
         read_size, insert_size_mean, insert_size_stddev, purity, with_normal_sample, preneoplastic_in_normal,
         include_non_sequenced_mutations, c_seed, write_SAM, output_dir, filename_prefix, template_name_prefix,
         samples=None, update_SAM=False, reference_genome=None):
    eng = pf.engine
    if not (0.0 <= float(purity) <= 1.0):
        raise ValueError("The purity must be in [0,1]")
    if insert_size_mean:
        from ._lib import check, lib
        check(lib().psx_insert_size_check(eng.ctx.h, int(insert_size_mean), int(insert_size_stddev)), eng.ctx.h)
    chrs = _relevant_chromosomes(eng.reference, chromosomes)
    seq = dict(coverage=float(coverage), read_size=int(read_size), insert_size_mean=int(insert_size_mean),
               insert_size_stddev=int(insert_size_stddev),
               error_rate=0.0 if sequencer is None else float(sequencer.error_rate), purity=float(purity),
               with_normal_sample=bool(with_normal_sample), preneoplastic_in_normal=bool(preneoplastic_in_normal),
               chromosomes=chrs, samples=samples)
    alt, depth, present, n_reads = core.simulate_seq_arrays(eng.ctx, eng.genome(), forest, pf.mutations,
                                                            _seed64(c_seed), **seq)
    names = list(sample_names) + (["normal_sample"] if with_normal_sample else [])
    if write_SAM:
        from .sam import write_sam
        write_sam(pf, forest, names, seq, _seed64(c_seed), output_dir, filename_prefix, template_name_prefix,
                  update=bool(update_SAM), genome=_read_genome(eng, reference_genome),
                  random_quality=bool(getattr(sequencer, "random_quality_scores", False)))
    return _table(pf, names, alt, depth, present, include_non_sequenced_mutations), n_reads


def _read_genome(eng: MutationEngine, reference_genome) -> core.Genome:
    """The genome reads are taken from: the engine's, or a reference_genome
    given to simulate_seq (src/seq_simulation.cpp:245-280), which must have
    the engine's chromosome names and lengths (the mutations' positions refer
    to them)."""
    if reference_genome is None:
        return eng.genome()
    if not isinstance(reference_genome, ReferenceGenome):
        raise ValueError("The reference genome must be a ReferenceGenome")
    if reference_genome.names != eng.reference.names or \
            not np.array_equal(reference_genome.lengths, eng.reference.lengths):
        raise ValueError("The reference genome must have the mutation engine's chromosome names and lengths")
    return core.Genome(eng.ctx, reference_genome.lengths, seed=reference_genome.seed, packed=reference_genome.packed)


def _check_sam_dir(write_SAM, update_SAM, output_dir):
    if write_SAM and not update_SAM and os.path.isdir(output_dir) and os.listdir(output_dir):
        raise ValueError(f"The output directory \"{output_dir}\" already exists")


def simulate_seq(phylo_forest: PhylogeneticForest, sequencer=None, reference_genome=None, chromosomes=None,
                 coverage=10, read_size=150, insert_size_mean=0, insert_size_stddev=10, output_dir="ProCESS_SAM",
                 write_SAM=False, update_SAM=False, cell_labelling: Callable | None = None, purity=1,
                 with_normal_sample=True, preneoplastic_in_normal=False, filename_prefix="chr_",
                 template_name_prefix="r", include_non_sequenced_mutations=False, seed=None) -> dict:
    """src/sequencing.cpp:193-209 and src/seq_simulation.cpp:517-601.
    reference_genome (a ReferenceGenome with the engine's chromosomes) only
    changes the bases of written reads; the counts do not read the
    reference."""
    c_seed = get_random_seed(seed)
    pf = phylo_forest
    if cell_labelling is not None:
        if not callable(cell_labelling):
            raise ValueError("The FACs_labelling_function must be a function.")
        forest, names = _labelled_forest(pf, cell_labelling)
    else:
        forest, names = pf.forest, list(pf.samples_forest.sample_names)
    _check_sam_dir(write_SAM, update_SAM, output_dir)
    df, _ = _run(pf, forest, names, sequencer, chromosomes, coverage, read_size, insert_size_mean,
                 insert_size_stddev, purity, with_normal_sample, preneoplastic_in_normal,
                 include_non_sequenced_mutations, c_seed, write_SAM, output_dir, filename_prefix,
                 template_name_prefix, update_SAM=update_SAM, reference_genome=reference_genome)
    parameters = {"sequencer": _sequencer_data(sequencer), "reference_genome": reference_genome,
                  "chromosomes": chromosomes, "coverage": coverage, "read_size": read_size,
                  "insert_size_mean": insert_size_mean, "insert_size_stddev": insert_size_stddev,
                  "output_dir": output_dir, "write_SAM": write_SAM, "update_SAM": update_SAM,
                  "cell_labelling": cell_labelling, "purity": purity, "with_normal_sample": with_normal_sample,
                  "filename_prefix": filename_prefix, "template_name_prefix": template_name_prefix,
                  "include_non_sequenced_mutations": include_non_sequenced_mutations, "seed": c_seed}
    return {"mutations": df, "parameters": parameters}


def simulate_normal_seq(phylo_forest: PhylogeneticForest, sequencer=None, reference_genome=None, chromosomes=None,
                        coverage=10, read_size=150, insert_size_mean=0, insert_size_stddev=10,
                        output_dir="ProCESS_normal_SAM", write_SAM=True, update_SAM=False, with_preneoplastic=False,
                        filename_prefix="chr_", template_name_prefix="r", include_non_sequenced_mutations=False,
                        seed=None) -> dict:
    """src/sequencing.cpp:268-282 and src/seq_simulation.cpp:603-679: the
    normal sample alone, purity 1."""
    c_seed = get_random_seed(seed)
    pf = phylo_forest
    _check_sam_dir(write_SAM, update_SAM, output_dir)
    df, _ = _run(pf, pf.forest, [], sequencer, chromosomes, coverage, read_size, insert_size_mean,
                 insert_size_stddev, 1.0, True, with_preneoplastic, include_non_sequenced_mutations, c_seed,
                 write_SAM, output_dir, filename_prefix, template_name_prefix, samples=[], update_SAM=update_SAM,
                 reference_genome=reference_genome)
    parameters = {"sequencer": _sequencer_data(sequencer), "reference_genome": reference_genome,
                  "chromosomes": chromosomes, "coverage": coverage, "read_size": read_size,
                  "insert_size_mean": insert_size_mean, "insert_size_stddev": insert_size_stddev,
                  "output_dir": output_dir, "write_SAM": write_SAM, "update_SAM": update_SAM,
                  "with_preneoplastic": with_preneoplastic, "filename_prefix": filename_prefix,
                  "template_name_prefix": template_name_prefix,
                  "include_non_sequenced_mutations": include_non_sequenced_mutations, "seed": c_seed}
    return {"mutations": df, "parameters": parameters}


def seq_to_long(seq_results) -> pd.DataFrame:
    """R/seq_to_long.R:31-67: wide per-sample columns to rows of
    chr, from, to, ref, alt, causes, classes, NV, DP, VAF, sample_name."""
    df = seq_results["mutations"] if isinstance(seq_results, dict) else seq_results
    names = [c[: -len(".VAF")] for c in df.columns if c.endswith(".VAF")]
    parts = []
    for sn in names:
        part = df[["chr", "chr_pos", "ref", "alt", "causes", "classes", f"{sn}.occurrences", f"{sn}.coverage",
                   f"{sn}.VAF"]].copy()
        part.columns = ["chr", "chr_pos", "ref", "alt", "causes", "classes", "NV", "DP", "VAF"]
        part["sample_name"] = sn
        parts.append(part)
    out = pd.concat(parts, ignore_index=True) if parts else pd.DataFrame()
    out = out.rename(columns={"chr_pos": "from"})
    out["to"] = out["from"]
    return out


def _wide(seq_results) -> pd.DataFrame:
    return seq_results["mutations"] if isinstance(seq_results, dict) else seq_results


def validate_chromosomes(seq_results, chromosomes=None) -> list:
    """R/ggplot_config.R:75-94: the table's chromosomes in order of first
    appearance when none are given, else the given ones, every one of which
    must occur in the table."""
    present = list(pd.unique(_wide(seq_results)["chr"]))
    if chromosomes is None:
        return present
    chromosomes = [chromosomes] if isinstance(chromosomes, str) else list(chromosomes)
    unknown = list(dict.fromkeys(c for c in chromosomes if c not in present))
    if unknown:
        names = ", ".join(str(c) for c in unknown)
        msg = f"The chromosomes {names} are" if len(unknown) > 1 else f"The chromosome {names} is"
        raise ValueError(msg + " not present in the sequence reference data.")
    return chromosomes


def get_seq_data(seq_results, sample, chromosomes=None) -> dict:
    """R/plot_genome_wide_mutations.R:3-19: the long table restricted to the
    chosen chromosomes (chr becomes a categorical in that order), split into
    the sample's rows ("tumour") and the normal sample's rows ("normal")."""
    chromosomes = validate_chromosomes(seq_results, chromosomes)
    data = seq_to_long(_wide(seq_results))
    data = data[data["chr"].isin(chromosomes)].copy()
    data["chr"] = pd.Categorical(data["chr"], categories=chromosomes)
    names = list(pd.unique(data["sample_name"]))
    if sample not in names:
        raise ValueError("Inserted 'sample' is not available, available samples are: " + ", ".join(names))
    return {"tumour": data[data["sample_name"] == sample].reset_index(drop=True),
            "normal": data[data["sample_name"] == "normal_sample"].reset_index(drop=True)}


def seq_dr(seq_results, sample, chromosomes=None) -> pd.DataFrame:
    """The data of R/plot_genome_wide_mutations.R:75-111 (plot_DR): the
    sample's rows left-joined with the normal sample's on (chr, from, ref,
    alt), DR = DP.tumour / DP.normal (inf or NaN where the normal depth is 0),
    ordered by (chromosome, position) with abs_pos = 1..n. plot_DR draws a
    random subset of N rows for display; every row is returned here."""
    df = _wide(seq_results)
    if not all(f"normal_sample.{k}" in df.columns for k in ("occurrences", "coverage", "VAF")):
        raise ValueError('The parameter "seq_result" does not contain the mandatory normal sample "normal_sample".')
    data = get_seq_data(df, sample, chromosomes)
    d = data["tumour"].merge(data["normal"], how="left", on=["chr", "from", "ref", "alt"],
                             suffixes=(".tumour", ".normal"))
    d["DR"] = d["DP.tumour"].astype(np.float64) / d["DP.normal"].astype(np.float64)
    d = d.sort_values(["chr", "from"], kind="mergesort").reset_index(drop=True)
    d["abs_pos"] = np.arange(1, len(d) + 1)
    return d


def seq_baf(seq_results, sample, chromosomes=None, cuts=(0.0, 1.0)) -> pd.DataFrame:
    """The data of R/plot_genome_wide_mutations.R:192-240 (plot_BAF): the
    sample's rows ordered by (chromosome, position), abs_pos = 1..n numbered
    before the VAF filter min(cuts) <= VAF <= max(cuts). plot_BAF draws a
    random subset of N rows for display; every row is returned here."""
    d = get_seq_data(seq_results, sample, chromosomes)["tumour"]
    d = d.sort_values(["chr", "from"], kind="mergesort").reset_index(drop=True)
    d["abs_pos"] = np.arange(1, len(d) + 1)
    lo, hi = min(cuts), max(cuts)
    return d[(d["VAF"] <= hi) & (d["VAF"] >= lo)].reset_index(drop=True)
