"""ctypes loader of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, always to check results against or as the CPU timing,
never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

from process_amd import _abi as A

HERE = Path(__file__).resolve().parent
# PSX_ORACLE_LIB names another build of the same sources (the sanitizer build of
# tests/test_oracle_sanitized.py); the default is the plain one
LIB_PATH = Path(os.environ["PSX_ORACLE_LIB"]) if os.environ.get("PSX_ORACLE_LIB") else HERE / "liboracle.so"
_lib = None


class OrcForest(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("n_leaves", C.c_uint32), ("n_samples", C.c_uint32),
                ("cell_id", A.i64p), ("parent", A.u32p), ("species", A.u32p), ("birth", A.f64p),
                ("leaf_lo", A.u32p), ("leaf_hi", A.u32p), ("sample_begin", A.u32p), ("sample_leaves", A.u32p)]


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        l = C.CDLL(str(LIB_PATH))
        vp, st = C.c_void_p, C.c_int
        fp = C.POINTER(C.POINTER(OrcForest))
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_free": (None, [vp]),
            "orc_set_threads": (C.c_int, [C.c_int]),
            "orc_genome_words": (C.c_uint64, [C.c_uint32, A.u32p]),
            "orc_genome_synth": (st, [C.c_uint32, A.u32p, C.c_uint64, A.u32p]),
            "orc_forest_flatten": (st, [C.POINTER(A.ForestDesc), fp]),
            "orc_forest_free": (None, [C.POINTER(OrcForest)]),
            "orc_place_mutations": (st, [C.c_uint32, A.u32p, A.u32p, C.POINTER(OrcForest),
                                         C.POINTER(A.EngineParams), C.c_uint64, C.POINTER(C.c_void_p), A.u64p]),
            "orc_sort_table": (st, [vp, C.c_uint64]),
            "orc_id_index": (st, [C.c_uint32, A.u32p, A.u32p, A.u64p, A.u8p, A.u64p, C.POINTER(C.c_void_p)]),
            "orc_sbs_index": (st, [C.c_uint32, A.u32p, A.u32p, C.c_uint32, A.u64p, A.u64p, C.POINTER(C.c_void_p)]),
            "orc_id_site": (C.c_int, [C.c_uint32, A.u32p, A.u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
            "orc_place_cna": (st, [C.c_uint32, A.u32p, C.POINTER(OrcForest), C.POINTER(A.EngineParams), C.c_uint64,
                                   C.POINTER(C.c_void_p), A.u32p]),
            "orc_simulate_seq_cna": (st, [C.c_uint32, A.u32p, C.POINTER(OrcForest), vp, C.c_uint64, C.c_uint32,
                                          vp, C.c_uint32, C.POINTER(A.SeqParams), C.c_uint64, A.i32p, A.i32p,
                                          A.u8p, A.u64p]),
            "orc_simulate_seq": (st, [C.c_uint32, A.u32p, C.POINTER(OrcForest), vp, C.c_uint64, C.c_uint32,
                                      C.POINTER(A.SeqParams), C.c_uint64, A.i32p, A.i32p, A.u8p, A.u64p]),
            "orc_simulate_reads": (st, [C.c_uint32, A.u32p, A.u32p, C.POINTER(OrcForest), vp, C.c_uint64,
                                        C.c_uint32, C.POINTER(A.SeqParams), C.c_uint64, A.u32p, C.c_uint64,
                                        A.u64p]),
            "orc_simulate_reads_cna": (st, [C.c_uint32, A.u32p, A.u32p, C.POINTER(OrcForest), vp, C.c_uint64,
                                            C.c_uint32, vp, C.c_uint32, C.POINTER(A.SeqParams), C.c_uint64, A.u32p,
                                            C.c_uint64, A.u64p]),
            "orc_refshape_seq": (st, [C.c_uint32, A.u32p, A.u32p, C.POINTER(OrcForest), vp, C.c_uint64, C.c_uint32,
                                      C.POINTER(A.SeqParams), C.c_uint64, C.c_int, A.i32p, A.i32p, A.u64p]),
            "orc_seq_table": (st, [vp, C.c_uint64, C.c_uint32, A.i32p, A.i32p, A.u8p, C.c_int32, A.u64p, A.u64p,
                                   A.i32p, A.i32p, A.f64p]),
            "orc_philox": (None, [A.u32p, C.c_uint64, A.u32p]),
            "orc_cdf_table": (st, [A.f64p, C.c_uint32, A.u64p]),
            "orc_poisson_table": (st, [C.c_double, A.u32p, A.u32p, A.u64p, C.c_uint32]),
            "orc_binomial_table": (st, [C.c_uint32, C.c_double, A.u32p, A.u32p, A.u64p, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


def set_threads(n: int) -> int:
    """OpenMP threads of simulate_seq (0 keeps the default); returns the count in use."""
    return int(lib().orc_set_threads(int(n)))


def philox(ctr, key: int) -> np.ndarray:
    c = np.ascontiguousarray(ctr, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox(A.ptr(c, C.c_uint32), C.c_uint64(key), A.ptr(out, C.c_uint32))
    return out


def cdf_table(pmf) -> np.ndarray:
    p = np.ascontiguousarray(pmf, np.float64)
    T = np.zeros(len(p), np.uint64)
    _check(lib().orc_cdf_table(A.ptr(p, C.c_double), len(p), A.ptr(T, C.c_uint64)))
    return T


def poisson_table(lam: float, cap: int = 1 << 16):
    T = np.zeros(cap, np.uint64)
    k0, n = C.c_uint32(), C.c_uint32()
    _check(lib().orc_poisson_table(lam, C.byref(k0), C.byref(n), A.ptr(T, C.c_uint64), cap))
    return k0.value, T[: n.value].copy()


def binomial_table(t: int, p: float, cap: int = 1 << 16):
    T = np.zeros(cap, np.uint64)
    k0, n = C.c_uint32(), C.c_uint32()
    _check(lib().orc_binomial_table(t, p, C.byref(k0), C.byref(n), A.ptr(T, C.c_uint64), cap))
    return k0.value, T[: n.value].copy()


def genome_synth(lens, seed: int) -> np.ndarray:
    L = np.ascontiguousarray(lens, np.uint32)
    out = np.zeros(int(lib().orc_genome_words(len(L), A.ptr(L, C.c_uint32))), np.uint32)
    _check(lib().orc_genome_synth(len(L), A.ptr(L, C.c_uint32), C.c_uint64(seed), A.ptr(out, C.c_uint32)))
    return out


class Forest:
    """Flattened forest owned by the oracle library."""

    def __init__(self, cell_id, parent, species, sample, birth, n_samples):
        desc, self._keep = A.forest_desc(cell_id, parent, species, sample, birth, n_samples)
        p = C.POINTER(OrcForest)()
        _check(lib().orc_forest_flatten(C.byref(desc), C.byref(p)))
        self.p = p
        f = p.contents
        self.n_nodes, self.n_leaves, self.n_samples = f.n_nodes, f.n_leaves, f.n_samples
        n = f.n_nodes
        self.cell_id = np.ctypeslib.as_array(f.cell_id, (n,)).copy()
        self.parent = np.ctypeslib.as_array(f.parent, (n,)).copy()
        self.species = np.ctypeslib.as_array(f.species, (n,)).copy()
        self.leaf_lo = np.ctypeslib.as_array(f.leaf_lo, (n,)).copy()
        self.leaf_hi = np.ctypeslib.as_array(f.leaf_hi, (n,)).copy()
        self.sample_begin = np.ctypeslib.as_array(f.sample_begin, (f.n_samples + 1,)).copy()
        ns = int(self.sample_begin[-1])
        self.sample_leaves = (np.ctypeslib.as_array(f.sample_leaves, (ns,)).copy() if ns else
                              np.zeros(0, np.uint32))

    def __del__(self):
        if getattr(self, "p", None) and _lib is not None:
            _lib.orc_forest_free(self.p)
            self.p = None


def place_mutations(lens, packed, forest: Forest, seed: int, **engine) -> np.ndarray:
    L = np.ascontiguousarray(lens, np.uint32)
    pk = np.ascontiguousarray(packed, np.uint32)
    params, keep = A.engine_params(**engine)
    out = C.c_void_p()
    n = C.c_uint64()
    _check(lib().orc_place_mutations(len(L), A.ptr(L, C.c_uint32), A.ptr(pk, C.c_uint32), forest.p,
                                     C.byref(params), C.c_uint64(seed), C.byref(out), C.byref(n)))
    nb = n.value * A.MUTATION_DTYPE.itemsize
    arr = np.frombuffer(C.string_at(out, nb), dtype=A.MUTATION_DTYPE).copy() if n.value else \
        np.zeros(0, A.MUTATION_DTYPE)
    lib().orc_free(out)
    del keep
    return arr


def place_cna(lens, forest: Forest, seed: int = 0, **engine) -> np.ndarray:
    """Copy-number drivers and passenger CNAs placed on the forest
    (CNA_EVENT_DTYPE, DFS order); seed: the placement seed (passengers)."""
    L = np.ascontiguousarray(lens, np.uint32)
    params, keep = A.engine_params(**engine)
    out = C.c_void_p()
    n = C.c_uint32()
    _check(lib().orc_place_cna(len(L), A.ptr(L, C.c_uint32), forest.p, C.byref(params), C.c_uint64(seed),
                               C.byref(out), C.byref(n)))
    nb = n.value * A.CNA_EVENT_DTYPE.itemsize
    arr = np.frombuffer(C.string_at(out, nb), dtype=A.CNA_EVENT_DTYPE).copy() if n.value else \
        np.zeros(0, A.CNA_EVENT_DTYPE)
    lib().orc_free(out)
    del keep
    return arr


def sbs_index(lens, packed, sampling: int = 1) -> dict:
    """SBS context index of a reference: counts[32] sites and kept[32]
    indexed sites per pyrimidine-centred context, and the list (global
    positions, ascending per context, contexts in order)."""
    L = np.ascontiguousarray(lens, np.uint32)
    pk = np.ascontiguousarray(packed, np.uint32)
    counts, kept = np.zeros(32, np.uint64), np.zeros(32, np.uint64)
    out = C.c_void_p()
    _check(lib().orc_sbs_index(len(L), A.ptr(L, C.c_uint32), A.ptr(pk, C.c_uint32), int(sampling),
                               A.ptr(counts, C.c_uint64), A.ptr(kept, C.c_uint64), C.byref(out)))
    n = int(kept.sum())
    lst = np.frombuffer(C.string_at(out, 4 * n), np.uint32).copy() if n else np.zeros(0, np.uint32)
    lib().orc_free(out)
    return {"counts": counts, "kept": kept, "list": lst}


def id_index(lens, packed) -> dict:
    """Indel (ID83) classes of a reference: counts[83] sites per class,
    listed[83] (sites listed), off[84] list ranges and the lists (global
    position << 10 | payload, ascending per class)."""
    L = np.ascontiguousarray(lens, np.uint32)
    pk = np.ascontiguousarray(packed, np.uint32)
    counts = np.zeros(A.PSX_ID_CLASSES, np.uint64)
    listed = np.zeros(A.PSX_ID_CLASSES, np.uint8)
    off = np.zeros(A.PSX_ID_CLASSES + 1, np.uint64)
    out = C.c_void_p()
    _check(lib().orc_id_index(len(L), A.ptr(L, C.c_uint32), A.ptr(pk, C.c_uint32), A.ptr(counts, C.c_uint64),
                              A.ptr(listed, C.c_uint8), A.ptr(off, C.c_uint64), C.byref(out)))
    n = int(off[-1])
    lst = np.frombuffer(C.string_at(out, 8 * n), np.uint64).copy() if n else np.zeros(0, np.uint64)
    lib().orc_free(out)
    return dict(counts=counts, listed=listed.astype(bool), off=off, list=lst)


def id_site(lens, packed, chr_: int, q: int, payload: int, cls: int) -> bool:
    L = np.ascontiguousarray(lens, np.uint32)
    pk = np.ascontiguousarray(packed, np.uint32)
    return bool(lib().orc_id_site(len(L), A.ptr(L, C.c_uint32), A.ptr(pk, C.c_uint32), chr_, q, payload, cls))


def sort_table(recs: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(recs, A.MUTATION_DTYPE).copy()
    _check(lib().orc_sort_table(r.ctypes.data_as(C.c_void_p), len(r)))
    return r


def simulate_seq(lens, forest: Forest, recs: np.ndarray, seed: int, ploidy: int = 2, cna=None, **seq):
    L = np.ascontiguousarray(lens, np.uint32)
    r = np.ascontiguousarray(recs, A.MUTATION_DTYPE)
    params, keep = A.seq_params(**seq)
    n_req = len(seq["samples"]) if seq.get("samples") is not None else forest.n_samples
    S = n_req + (1 if seq.get("with_normal_sample", True) else 0)
    M = len(r)
    alt = np.zeros(S * M, np.int32)
    dep = np.zeros(S * M, np.int32)
    pres = np.zeros(S * M, np.uint8)
    nr = C.c_uint64()
    ev = np.ascontiguousarray(cna if cna is not None else np.zeros(0, A.CNA_EVENT_DTYPE), A.CNA_EVENT_DTYPE)
    _check(lib().orc_simulate_seq_cna(len(L), A.ptr(L, C.c_uint32), forest.p, r.ctypes.data_as(C.c_void_p), M,
                                      ploidy, ev.ctypes.data_as(C.c_void_p) if len(ev) else None, len(ev),
                                      C.byref(params), C.c_uint64(seed), A.ptr(alt, C.c_int32),
                                      A.ptr(dep, C.c_int32), A.ptr(pres, C.c_uint8), C.byref(nr)))
    del keep
    return alt.reshape(S, M), dep.reshape(S, M), pres.reshape(S, M), nr.value


def simulate_reads(lens, packed, forest: Forest, recs: np.ndarray, seed: int, capacity: int, ploidy: int = 2,
                   cna=None, **seq) -> np.ndarray:
    """Every read of simulate_seq, materialised: (n_reads, 2 + ceil(L / 16))
    uint32 rows in the layout and order of psx_simulate_seq_reads; cna:
    copy-number events (CNA_EVENT_DTYPE, DFS order) or None."""
    L = np.ascontiguousarray(lens, np.uint32)
    P = np.ascontiguousarray(packed, np.uint32)
    r = np.ascontiguousarray(recs, A.MUTATION_DTYPE)
    ev = np.ascontiguousarray(cna if cna is not None else np.zeros(0, A.CNA_EVENT_DTYPE), A.CNA_EVENT_DTYPE)
    params, keep = A.seq_params(**seq)
    rw = 2 + (int(seq["read_size"]) + 15) // 16
    rows = np.zeros(max(capacity, 1) * rw, np.uint32)
    nr = C.c_uint64()
    _check(lib().orc_simulate_reads_cna(len(L), A.ptr(L, C.c_uint32), A.ptr(P, C.c_uint32), forest.p,
                                        r.ctypes.data_as(C.c_void_p), len(r), ploidy,
                                        ev.ctypes.data_as(C.c_void_p) if len(ev) else None, len(ev),
                                        C.byref(params), C.c_uint64(seed), A.ptr(rows, C.c_uint32), capacity,
                                        C.byref(nr)))
    del keep
    return rows[: nr.value * rw].reshape(nr.value, rw)


def refshape_seq(lens, packed, forest: Forest, recs: np.ndarray, seed: int, threads: int = 1, ploidy: int = 2,
                 **seq):
    """The reference-shaped restatement (per-read base and quality strings,
    ordered-map lookups and statistics): (alt, depth, n_reads) equal to
    simulate_seq's; threads = OpenMP threads."""
    L = np.ascontiguousarray(lens, np.uint32)
    P = np.ascontiguousarray(packed, np.uint32)
    r = np.ascontiguousarray(recs, A.MUTATION_DTYPE)
    params, keep = A.seq_params(**seq)
    n_req = len(seq["samples"]) if seq.get("samples") is not None else forest.n_samples
    S = n_req + (1 if seq.get("with_normal_sample", True) else 0)
    M = len(r)
    alt = np.zeros(S * max(M, 1), np.int32)
    dep = np.zeros(S * max(M, 1), np.int32)
    nr = C.c_uint64()
    _check(lib().orc_refshape_seq(len(L), A.ptr(L, C.c_uint32), A.ptr(P, C.c_uint32), forest.p,
                                  r.ctypes.data_as(C.c_void_p), M, ploidy, C.byref(params), C.c_uint64(seed),
                                  int(threads), A.ptr(alt, C.c_int32), A.ptr(dep, C.c_int32), C.byref(nr)))
    del keep
    return alt[: S * M].reshape(S, M), dep[: S * M].reshape(S, M), nr.value


def seq_table(recs, alt, depth, present, include_non_sequenced=False):
    r = np.ascontiguousarray(recs, A.MUTATION_DTYPE)
    S, M = alt.shape
    a = np.ascontiguousarray(alt, np.int32)
    d = np.ascontiguousarray(depth, np.int32)
    p = np.ascontiguousarray(present, np.uint8)
    rows = np.zeros(max(M, 1), np.uint64)
    occ = np.zeros(S * max(M, 1), np.int32)
    cov = np.zeros(S * max(M, 1), np.int32)
    vaf = np.zeros(S * max(M, 1), np.float64)
    nr = C.c_uint64()
    _check(lib().orc_seq_table(r.ctypes.data_as(C.c_void_p), M, S, A.ptr(a, C.c_int32), A.ptr(d, C.c_int32),
                               A.ptr(p, C.c_uint8), int(bool(include_non_sequenced)), C.byref(nr),
                               A.ptr(rows, C.c_uint64), A.ptr(occ, C.c_int32), A.ptr(cov, C.c_int32),
                               A.ptr(vaf, C.c_double)))
    n = nr.value
    Mx = max(M, 1)
    return (rows[:n].copy(), occ.reshape(S, Mx)[:, :n].copy(), cov.reshape(S, Mx)[:, :n].copy(),
            vaf.reshape(S, Mx)[:, :n].copy())
