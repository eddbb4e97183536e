"""The SBS context index of the oracle (the reference engine's context index,
src/mutation_engine.cpp:213-261, with context_sampling and the driver loci
left out): the index against a brute-force enumeration over the bases as
strings (reverse complement for a purine centre), the sampling law (every
s-th site of a context, from the first), and placement on crafted references
where the answer is fixed: every SNV sits on an indexed site of its class,
never on a driver locus, spread uniformly over the class's sites, and classes
whose context has no indexed site are left out of the signature."""
from __future__ import annotations

import numpy as np
import pytest
from scipy import stats

import cases
from oracle import oracle as O
from process_amd import _abi as A

BASE = "ACGT"
COMP = {"A": "T", "C": "G", "G": "C", "T": "A"}


def pack(seqs):
    words = []
    for s in seqs:
        b = np.array([BASE.index(ch) for ch in s], np.uint32)
        b = np.concatenate([b, np.zeros((-len(b)) % 16, np.uint32)]).reshape(-1, 16)
        words.append((b << (2 * np.arange(16, dtype=np.uint32))).sum(axis=1, dtype=np.uint64).astype(np.uint32))
    return np.concatenate(words), np.array([len(s) for s in seqs], np.uint32)


def context_of(tri: str) -> int:
    """Context code of a trinucleotide: read on the strand with a pyrimidine
    centre, 5' base << 2 | 3' base, + 16 for a T centre."""
    if tri[1] in "AG":
        tri = "".join(COMP[ch] for ch in reversed(tri))
    return (16 if tri[1] == "T" else 0) | (BASE.index(tri[0]) << 2) | BASE.index(tri[2])


def brute_lists(seqs):
    lists = [[] for _ in range(32)]
    g0 = 0
    for s in seqs:
        for q in range(1, len(s) - 1):
            lists[context_of(s[q - 1:q + 2])].append(g0 + q)
        g0 += len(s)
    return lists


def random_seqs(seed, lens):
    rng = np.random.default_rng(seed)
    return ["".join(rng.choice(list(BASE), n)) for n in lens]


@pytest.mark.parametrize("sampling", [1, 3, 100])
def test_index_matches_brute_force(sampling):
    seqs = random_seqs(5, [4000, 2999, 3, 2, 1500])
    packed, lens = pack(seqs)
    ix = O.sbs_index(lens, packed, sampling)
    full = brute_lists(seqs)
    assert ix["counts"].tolist() == [len(l) for l in full]
    assert int(ix["counts"].sum()) == sum(max(0, n - 2) for n in lens)
    kept = [l[::sampling] for l in full]
    assert ix["kept"].tolist() == [len(l) for l in kept]
    assert ix["list"].tolist() == [x for l in kept for x in l]


def test_context_codes_are_the_classes_contexts():
    """Class cls (substitution sub = cls // 16 in C>A C>G C>T T>A T>C T>G,
    5' base (cls // 4) % 4, 3' base cls % 4) samples the context of its
    pyrimidine trinucleotide."""
    for cls in range(96):
        sub, five, three = cls // 16, (cls // 4) % 4, cls % 4
        tri = BASE[five] + ("C" if sub < 3 else "T") + BASE[three]
        assert context_of(tri) == ((16 if sub >= 3 else 0) | (cls & 15))
        # the purine-centred reverse complement is the same context
        assert context_of("".join(COMP[ch] for ch in reversed(tri))) == context_of(tri)


def forest():
    fd = cases.forest(n_cells=600, n_samples=1, sample_size=200, seed=7, two_species=False)
    return O.Forest(fd["cell_id"], fd["parent"], fd["species"], fd["sample"], fd["birth"], fd["n_samples"])


def one_class_engine(cls, *, snv, drivers=(), pre=0, infinite_sites=False, sampling=1, extra=None):
    sig = np.zeros(96)
    sig[cls] = 1.0
    if extra is not None:
        sig[extra] = 1.0
    sig /= sig.sum()
    return dict(snv_rate=[snv], indel_rate=[0.0], sbs=sig[None, :], epoch_start=[0.0], exposure=[[1.0]],
                n_pre_snv=pre, n_pre_indel=0, pre_sbs=0, indel_max_len=10, indel_len_p=0.5,
                infinite_sites=infinite_sites, ploidy=2, drivers=list(drivers), germline=[],
                context_sampling=sampling)


ACA = 2 * 16 + (0 << 2) | 0  # A[C>T]A


def test_snvs_sit_on_indexed_sites_of_their_class():
    seqs = random_seqs(9, [30_000, 17_000])
    packed, lens = pack(seqs)
    sig = cases.signatures(2, 3)
    eng = dict(snv_rate=[4e-6], indel_rate=[0.0], sbs=sig, epoch_start=[0.0, 4.0],
               exposure=[[0.5, 0.5], [0.1, 0.9]], n_pre_snv=30, n_pre_indel=0, pre_sbs=1, indel_max_len=10,
               indel_len_p=0.5, infinite_sites=True, ploidy=2, drivers=[], germline=[], context_sampling=7)
    m = O.place_mutations(lens, packed, forest(), 11, **eng)
    snv = m[m["kind"] == A.PSX_SNV]
    assert len(snv) > 200
    ix = O.sbs_index(lens, packed, 7)
    off = np.concatenate([[0], np.cumsum(ix["kept"])]).astype(np.int64)
    boff = np.concatenate([[0], np.cumsum(lens.astype(np.int64))])
    for r in snv:
        q = int(r["pos"])
        tri = seqs[int(r["chr"])][q - 1:q + 2]
        # the record's reference base and the class's context
        assert BASE.index(tri[1]) == int(r["ref"])
        ctx = context_of(tri)
        assert int(boff[int(r["chr"])]) + q in set(ix["list"][int(off[ctx]):int(off[ctx + 1])].tolist())
        # the alternative is the class's substitution on the site's strand
        alt = BASE[int(r["alt"])]
        pyr_alt = alt if tri[1] in "CT" else COMP[alt]
        assert pyr_alt != ("C" if ctx < 16 else "T")


def crafted(n_sites, spacing=97):
    """Poly-A with n_sites isolated ACA trinucleotides: context A[C]A has
    exactly those sites, T[T]T (the reverse complement of AAA) all others."""
    s = ["A"] * (n_sites * spacing + 50)
    sites = [20 + k * spacing for k in range(n_sites)]
    for q in sites:
        s[q] = "C"
    return "".join(s), sites


def test_driver_loci_are_never_drawn():
    seq, sites = crafted(2)
    packed, lens = pack([seq])
    drv = [(0, 0, sites[0], 2, -1)]  # a C>G driver on the first ACA
    eng = one_class_engine(ACA, snv=6e-4, drivers=drv, pre=5)
    m = O.place_mutations(lens, packed, forest(), 3, **eng)
    pas = m[(m["kind"] == A.PSX_SNV) & (m["nature"] != A.PSX_DRIVER)]
    assert len(pas) > 100
    assert set(pas["pos"].tolist()) == {sites[1]} and set(pas["alt"].tolist()) == {3}
    assert (m[m["nature"] == A.PSX_DRIVER]["pos"] == sites[0]).all()


def test_sites_uniform_over_the_context():
    seq, sites = crafted(12)
    packed, lens = pack([seq])
    eng = one_class_engine(ACA, snv=8e-4)
    m = O.place_mutations(lens, packed, forest(), 8, **eng)
    pos = m[m["kind"] == A.PSX_SNV]["pos"]
    counts = np.array([(pos == q).sum() for q in sites])
    assert counts.sum() == len(pos) > 1000
    assert stats.chisquare(counts).pvalue > 1e-3
    # sampling 4 keeps sites 0, 4, 8 of the context
    m4 = O.place_mutations(lens, packed, forest(), 8, **dict(eng, context_sampling=4))
    assert set(m4[m4["kind"] == A.PSX_SNV]["pos"].tolist()) == {sites[0], sites[4], sites[8]}


def test_classes_without_indexed_sites_are_left_out():
    seq, sites = crafted(3)
    packed, lens = pack([seq])
    # half the mass on A[C>T]A, half on C[C>T]G, which the reference lacks
    eng = one_class_engine(ACA, snv=4e-4, extra=2 * 16 + (1 << 2) | 2)
    m = O.place_mutations(lens, packed, forest(), 4, **eng)
    snv = m[m["kind"] == A.PSX_SNV]
    assert len(snv) > 100 and set(snv["pos"].tolist()) <= set(sites)
    with pytest.raises(O.OracleError, match="no class with a site in the context index"):
        O.place_mutations(lens, packed, forest(), 4, **one_class_engine(2 * 16 + (1 << 2) | 2, snv=2e-4))
