"""ctypes mirror of include/psx.h (struct layouts and the numpy record dtype)."""
from __future__ import annotations

import ctypes as C

import numpy as np

PSX_STRATUM = 8192
PSX_ERR_KCAP = 128
PSX_MAX_CHROMOSOMES = 64
PSX_NO_NODE = 0xFFFFFFFF
PSX_NO_CAUSE = 0xFFFF

PSX_SNV, PSX_DEL, PSX_INS = 0, 1, 2
PSX_PASSENGER, PSX_DRIVER, PSX_PRENEOPLASTIC, PSX_GERMINAL = 0, 1, 2, 3

STATUS = {0: "PSX_OK", 1: "PSX_EINVAL", 2: "PSX_ENOMEM_DEVICE", 3: "PSX_EHIP",
          4: "PSX_ECANCELLED", 5: "PSX_ENOMEM_HOST", 6: "PSX_EUNSUPPORTED"}

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)


class ForestDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("cell_id", i64p), ("parent", i64p), ("species", u32p),
                ("sample", i32p), ("birth_time", f64p), ("n_samples", C.c_uint32)]


class Driver(C.Structure):
    _fields_ = [("species", C.c_uint32), ("chr", C.c_uint32), ("pos", C.c_uint32), ("alt", C.c_uint8),
                ("allele", C.c_int8), ("pad", C.c_uint8 * 2)]


class Germline(C.Structure):
    _fields_ = [("chr", C.c_uint32), ("pos", C.c_uint32), ("alt", C.c_uint8), ("allele", C.c_uint8),
                ("pad", C.c_uint8 * 2)]


PSX_CNA_AMP, PSX_CNA_DEL, PSX_CNA_WGD = 0, 1, 2
PSX_ALL_CHROMOSOMES = 0xFFFFFFFF
PSX_CNA_RANDOM_ALLELE = 0xFF
PSX_PROBE_PHILOX, PSX_PROBE_INT = 0, 1
PSX_PROBE_STORE_FILL, PSX_PROBE_STORE_CHUNKS, PSX_PROBE_STORE_CHUNKS_SHIFTED = 0, 1, 2


class Cna(C.Structure):
    _fields_ = [("species", C.c_uint32), ("chr", C.c_uint32), ("start", C.c_uint32), ("len", C.c_uint32),
                ("allele", C.c_uint8), ("kind", C.c_uint8), ("copies", C.c_uint8), ("pad", C.c_uint8)]


class CnaRegion(C.Structure):
    _fields_ = [("chr", C.c_uint32), ("start", C.c_uint32), ("len", C.c_uint32), ("kind", C.c_uint8),
                ("pad", C.c_uint8 * 3)]


PSX_ID_CLASSES = 83

CNA_EVENT_DTYPE = np.dtype([("node", "<u4"), ("leaf_lo", "<u4"), ("leaf_hi", "<u4"), ("chr", "<u4"),
                            ("start", "<u4"), ("end", "<u4"), ("allele", "u1"), ("kind", "u1"), ("copies", "u1"),
                            ("pad", "u1")])


class EngineParams(C.Structure):
    _fields_ = [("n_species", C.c_uint32), ("snv_rate", f64p), ("indel_rate", f64p),
                ("n_sbs", C.c_uint32), ("sbs", f64p), ("n_epochs", C.c_uint32), ("epoch_start", f64p),
                ("exposure", f64p), ("n_preneoplastic_snv", C.c_uint32),
                ("n_preneoplastic_indel", C.c_uint32), ("preneoplastic_sbs", C.c_uint32),
                ("indel_max_len", C.c_uint32), ("indel_len_p", C.c_double), ("infinite_sites", C.c_int32),
                ("ploidy", C.c_uint32), ("n_drivers", C.c_uint32), ("drivers", C.POINTER(Driver)),
                ("n_germline", C.c_uint32), ("germline", C.POINTER(Germline)),
                ("n_cna", C.c_uint32), ("cna", C.POINTER(Cna)),
                ("cna_rate", f64p), ("n_cna_catalogue", C.c_uint32), ("cna_catalogue", C.POINTER(CnaRegion)),
                ("n_id", C.c_uint32), ("id", f64p), ("id_exposure", f64p), ("preneoplastic_id", C.c_uint32),
                ("context_sampling", C.c_uint32)]


class SeqParams(C.Structure):
    _fields_ = [("coverage", C.c_double), ("read_size", C.c_uint32), ("insert_size_mean", C.c_uint32),
                ("insert_size_stddev", C.c_uint32), ("error_rate", C.c_double), ("purity", C.c_double),
                ("with_normal_sample", C.c_int32), ("preneoplastic_in_normal", C.c_int32),
                ("n_chromosomes", C.c_uint32), ("chromosomes", u32p), ("n_samples", C.c_uint32),
                ("samples", u32p)]


class SeqJob(C.Structure):
    _fields_ = [("genome", C.c_void_p), ("forest", C.c_void_p), ("mutations", C.c_void_p),
                ("params", SeqParams), ("seed", C.c_uint64), ("d_alt", C.c_void_p), ("d_depth", C.c_void_p),
                ("d_present", C.c_void_p)]


class SamParams(C.Structure):
    _fields_ = [("output_dir", C.c_char_p), ("filename_prefix", C.c_char_p), ("template_name_prefix", C.c_char_p),
                ("update", C.c_int32), ("random_quality", C.c_int32), ("sample_names", C.POINTER(C.c_char_p)),
                ("chr_names", C.POINTER(C.c_char_p))]


class LaunchConfig(C.Structure):
    _fields_ = [("block_threads", C.c_uint32), ("lds_records", C.c_uint32), ("strata_per_item", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("use_graph", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_inflight", C.c_uint32), ("max_batch", C.c_uint32),
                ("launch", LaunchConfig), ("overlap_after_us", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("genome", C.c_void_p), ("forest", C.c_void_p), ("mutations", C.c_void_p),
                ("params", SeqParams), ("seed", C.c_uint64), ("alt", i32p), ("depth", i32p), ("present", u8p)]


class RequestStats(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_reads", C.c_uint64), ("batch_size", C.c_uint32), ("lane", C.c_uint32),
                ("queued_ms", C.c_double), ("latency_ms", C.c_double)]


class ServerTotals(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("launches", C.c_uint64), ("reads", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("peak_bytes", C.c_uint64)]


class SynthParams(C.Structure):
    _fields_ = [("n_species", C.c_uint32), ("growth", f64p), ("death", f64p), ("intro_at", u64p),
                ("n_cells", C.c_uint64), ("n_samples", C.c_uint32), ("sample_size", C.c_uint64)]


MUTATION_DTYPE = np.dtype({
    "names": ["chr", "pos", "leaf_lo", "leaf_hi", "node", "seq", "cause", "allele", "kind", "ref", "alt",
              "len", "nature", "sid_head", "pad"],
    "formats": ["<u4", "<u4", "<u4", "<u4", "<u4", "<u4", "<u2", "u1", "u1", "u1", "u1", "u1", "u1", "u1",
                ("u1", (3,))],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 26, 27, 28, 29, 30, 31, 32, 33],
    "itemsize": 36,
})


def ptr(a: np.ndarray | None, ctype):
    """Pointer to a contiguous numpy array (None -> NULL)."""
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def forest_desc(cell_id, parent, species, sample, birth, n_samples) -> tuple[ForestDesc, list]:
    keep = [np.ascontiguousarray(cell_id, np.int64), np.ascontiguousarray(parent, np.int64),
            np.ascontiguousarray(species, np.uint32), np.ascontiguousarray(sample, np.int32),
            np.ascontiguousarray(birth, np.float64)]
    d = ForestDesc(len(keep[0]), ptr(keep[0], C.c_int64), ptr(keep[1], C.c_int64), ptr(keep[2], C.c_uint32),
                   ptr(keep[3], C.c_int32), ptr(keep[4], C.c_double), int(n_samples))
    return d, keep


def engine_params(*, snv_rate, indel_rate, sbs, epoch_start, exposure, n_pre_snv=0, n_pre_indel=0,
                  pre_sbs=0, indel_max_len=10, indel_len_p=0.5, infinite_sites=True, ploidy=2,
                  drivers=(), germline=(), cna=(), cna_rate=None, cna_catalogue=(), id_sig=None,
                  id_exposure=None, pre_id=0, context_sampling=1) -> tuple[EngineParams, list]:
    """cna: (species, chr, start, len, allele, kind, copies) tuples (kind PSX_CNA_*); cna_rate: passenger
    CNAs per base per division per species (None: none); cna_catalogue: (chr, start, len, kind) regions;
    id_sig: [n_id][83] indel signatures; id_exposure: [n_epochs][n_id]; context_sampling: every how many
    sites of a trinucleotide context the SBS context index keeps (1: all)."""
    snv = np.ascontiguousarray(snv_rate, np.float64)
    ind = np.ascontiguousarray(indel_rate, np.float64)
    sbs_a = np.ascontiguousarray(sbs, np.float64).reshape(-1, 96)
    es = np.ascontiguousarray(epoch_start, np.float64).reshape(-1)
    ex = np.ascontiguousarray(exposure, np.float64).reshape(len(es), -1) if len(es) else np.zeros((0, 1))
    drv = (Driver * max(1, len(drivers)))()
    for i, d in enumerate(drivers):
        drv[i] = Driver(int(d[0]), int(d[1]), int(d[2]), int(d[3]), int(d[4]) if len(d) > 4 else -1)
    grm = (Germline * max(1, len(germline)))()
    for i, g in enumerate(germline):
        grm[i] = Germline(int(g[0]), int(g[1]), int(g[2]), int(g[3]))
    cn = (Cna * max(1, len(cna)))()
    for i, c in enumerate(cna):
        cn[i] = Cna(*(int(v) for v in c[:7]))
    crate = None if cna_rate is None else np.ascontiguousarray(cna_rate, np.float64)
    cat = (CnaRegion * max(1, len(cna_catalogue)))()
    for i, r in enumerate(cna_catalogue):
        cat[i] = CnaRegion(int(r[0]), int(r[1]), int(r[2]), int(r[3]))
    ids = None if id_sig is None else np.ascontiguousarray(id_sig, np.float64).reshape(-1, PSX_ID_CLASSES)
    n_id = 0 if ids is None else ids.shape[0]
    idx = None if ids is None else np.ascontiguousarray(id_exposure, np.float64).reshape(len(es), n_id)
    p = EngineParams(len(snv), ptr(snv, C.c_double), ptr(ind, C.c_double), sbs_a.shape[0], ptr(sbs_a, C.c_double),
                     len(es), ptr(es, C.c_double), ptr(np.ascontiguousarray(ex), C.c_double), int(n_pre_snv),
                     int(n_pre_indel), int(pre_sbs), int(indel_max_len), float(indel_len_p), int(bool(infinite_sites)),
                     int(ploidy), len(drivers), C.cast(drv, C.POINTER(Driver)), len(germline),
                     C.cast(grm, C.POINTER(Germline)), len(cna), C.cast(cn, C.POINTER(Cna)),
                     ptr(crate, C.c_double), len(cna_catalogue), C.cast(cat, C.POINTER(CnaRegion)),
                     n_id, ptr(ids, C.c_double), ptr(idx, C.c_double), int(pre_id), int(context_sampling))
    return p, [snv, ind, sbs_a, es, ex, drv, grm, cn, crate, cat, ids, idx]


def seq_params(*, coverage=10.0, read_size=150, insert_size_mean=0, insert_size_stddev=10, error_rate=0.0,
               purity=1.0, with_normal_sample=True, preneoplastic_in_normal=False, chromosomes=None,
               samples=None) -> tuple[SeqParams, list]:
    ch = None if chromosomes is None else np.ascontiguousarray(chromosomes, np.uint32)
    # an empty list still passes a non-NULL pointer: no tumour sample
    sm = None if samples is None else np.ascontiguousarray(samples, np.uint32) if len(samples) else \
        np.zeros(1, np.uint32)
    p = SeqParams(float(coverage), int(read_size), int(insert_size_mean), int(insert_size_stddev),
                  float(error_rate), float(purity), int(bool(with_normal_sample)), int(bool(preneoplastic_in_normal)),
                  0 if ch is None else len(ch), ptr(ch, C.c_uint32), 0 if samples is None else len(samples),
                  ptr(sm, C.c_uint32))
    return p, [ch, sm]
