"""Thin owners of the libpsx device handles (context, genome, forest, mutation
table) plus the array-level calls of the C ABI. The R-style API in
process_amd.api is built on these."""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from . import _abi as A
from ._lib import PsxError, check, lib

_default_ctx = None
_ctx_lock = threading.Lock()


class Context:
    """One HIP stream on one device (psx_ctx)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        st = lib().psx_ctx_create(device, C.byref(h))
        self.h = h
        if st != 0:
            try:
                check(st, h)
            finally:
                lib().psx_ctx_destroy(h)
                self.h = None
        self.device = device

    def close(self):
        if self.h:
            lib().psx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        check(lib().psx_ctx_synchronize(self.h), self.h)

    def last_kernel_ms(self) -> float:
        v = C.c_double()
        check(lib().psx_ctx_last_kernel_ms(self.h, C.byref(v)), self.h)
        return v.value

    def last_place_ms(self) -> float:
        v = C.c_double()
        check(lib().psx_ctx_last_place_ms(self.h, C.byref(v)), self.h)
        return v.value

    def cancel(self) -> None:
        """Stops the call running on this context (from another thread) at its
        next work-chunk boundary with PsxError(PSX_ECANCELLED) (psx_ctx_cancel)."""
        check(lib().psx_ctx_cancel(self.h), self.h)

    def mem_info(self) -> tuple[int, int]:
        f, t = C.c_uint64(), C.c_uint64()
        check(lib().psx_ctx_mem_info(self.h, C.byref(f), C.byref(t)), self.h)
        return f.value, t.value

    def trim(self) -> None:
        """Frees the buffers the context keeps grown between calls; later calls
        grow them again (psx_ctx_trim)."""
        check(lib().psx_ctx_trim(self.h), self.h)


class DeviceArray:
    """A device buffer owned through libpsx (psx_dev_alloc), with numpy copies.

    libpsx and the ROCm runtime bundled with torch are different HIP runtimes
    and cannot share the GPU in one process, so device buffers for the
    device-pointer entry points come from the library itself."""

    def __init__(self, ctx: Context, shape, dtype):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().psx_dev_alloc(ctx.h, self.nbytes, C.byref(p)), ctx.h)
        self.ptr = p.value

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().psx_dev_download(self.ctx.h, out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes), self.ctx.h)
        return out

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().psx_dev_upload(self.ctx.h, self.ptr, a.ctypes.data_as(C.c_void_p), self.nbytes), self.ctx.h)
        return self

    def fill(self, byte: int) -> "DeviceArray":
        check(lib().psx_dev_fill_u8(self.ctx.h, self.ptr, byte, self.nbytes), self.ctx.h)
        return self

    def close(self):
        if self.ptr:
            lib().psx_dev_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context() -> Context:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


class _Handle:
    _free = None

    def __init__(self, ctx: Context, h: C.c_void_p):
        self.ctx = ctx
        self.h = h

    def close(self):
        if self.h:
            getattr(lib(), self._free)(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Genome(_Handle):
    """Device-resident 2-bit packed reference (psx_genome)."""

    _free = "psx_genome_free"

    def __init__(self, ctx: Context, lengths, *, seed: int | None = None, packed=None):
        L = np.ascontiguousarray(lengths, np.uint32)
        h = C.c_void_p()
        if packed is None:
            check(lib().psx_genome_synth(ctx.h, len(L), A.ptr(L, C.c_uint32), C.c_uint64(seed or 0), C.byref(h)),
                  ctx.h)
        else:
            pk = np.ascontiguousarray(packed, np.uint32)
            check(lib().psx_genome_upload(ctx.h, len(L), A.ptr(L, C.c_uint32), A.ptr(pk, C.c_uint32), C.byref(h)),
                  ctx.h)
        super().__init__(ctx, h)
        self.lengths = L
        self.seed = seed

    def download(self) -> np.ndarray:
        n = int(lib().psx_genome_words(len(self.lengths), A.ptr(self.lengths, C.c_uint32)))
        out = np.zeros(n, np.uint32)
        check(lib().psx_genome_download(self.ctx.h, self.h, A.ptr(out, C.c_uint32), n), self.ctx.h)
        return out

    def context_index(self, sampling: int = 1, with_list: bool = False) -> dict:
        """The SBS context index at a sampling (built on the GPU on first
        use): counts[32] and kept[32] sites per context, build_ms, and with
        with_list the indexed global positions (ascending per context)."""
        counts, kept = np.zeros(32, np.uint64), np.zeros(32, np.uint64)
        ms = C.c_double()
        check(lib().psx_genome_context_index(self.ctx.h, self.h, int(sampling), A.ptr(counts, C.c_uint64),
                                             A.ptr(kept, C.c_uint64), None, 0, C.byref(ms)), self.ctx.h)
        out = {"counts": counts, "kept": kept, "build_ms": ms.value}
        if with_list:
            lst = np.zeros(max(1, int(kept.sum())), np.uint32)
            check(lib().psx_genome_context_index(self.ctx.h, self.h, int(sampling), None, None,
                                                 A.ptr(lst, C.c_uint32), len(lst), None), self.ctx.h)
            out["list"] = lst[:int(kept.sum())]
        return out


class Forest(_Handle):
    """Flattened, device-resident samples forest (psx_forest)."""

    _free = "psx_forest_free"

    def __init__(self, ctx: Context, cell_id, parent, species, sample, birth, n_samples: int):
        desc, keep = A.forest_desc(cell_id, parent, species, sample, birth, n_samples)
        h = C.c_void_p()
        check(lib().psx_forest_upload(ctx.h, C.byref(desc), C.byref(h)), ctx.h)
        del keep
        super().__init__(ctx, h)
        nn, nl, ns = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().psx_forest_info(h, C.byref(nn), C.byref(nl), C.byref(ns)))
        self.n_nodes, self.n_leaves, self.n_samples = nn.value, nl.value, ns.value

    def flat(self) -> dict:
        n, nl, ns = self.n_nodes, self.n_leaves, self.n_samples
        out = dict(cell_id=np.zeros(n, np.int64), parent=np.zeros(n, np.uint32), leaf_lo=np.zeros(n, np.uint32),
                   leaf_hi=np.zeros(n, np.uint32), sample_begin=np.zeros(ns + 1, np.uint32),
                   sample_leaves=np.zeros(max(nl, 1), np.uint32))
        check(lib().psx_forest_flat(self.h, A.ptr(out["cell_id"], C.c_int64), A.ptr(out["parent"], C.c_uint32),
                                    A.ptr(out["leaf_lo"], C.c_uint32), A.ptr(out["leaf_hi"], C.c_uint32),
                                    A.ptr(out["sample_begin"], C.c_uint32), A.ptr(out["sample_leaves"], C.c_uint32)))
        out["sample_leaves"] = out["sample_leaves"][: int(out["sample_begin"][-1])]
        return out


class Mutations(_Handle):
    """Device-resident mutation table in SID order (psx_mutations)."""

    _free = "psx_mutations_free"

    def __init__(self, ctx: Context, h: C.c_void_p, ploidy: int):
        super().__init__(ctx, h)
        n = C.c_uint64()
        check(lib().psx_mutations_count(h, C.byref(n)))
        self.n = n.value
        self.ploidy = ploidy

    @classmethod
    def place(cls, ctx: Context, genome: Genome, forest: Forest, seed: int, **engine) -> "Mutations":
        params, keep = A.engine_params(**engine)
        h = C.c_void_p()
        check(lib().psx_place_mutations(ctx.h, genome.h, forest.h, C.byref(params), C.c_uint64(seed), C.byref(h)),
              ctx.h)
        del keep
        return cls(ctx, h, int(engine.get("ploidy", 2)))

    @classmethod
    def place_batch(cls, ctx: Context, genome: Genome, forests: list[Forest], seeds, engines: list[dict]) \
            -> list["Mutations"]:
        """Placement of many forests in one call (psx_place_mutations_batch):
        table k equals Mutations.place(ctx, genome, forests[k], seeds[k],
        **engines[k])."""
        n = len(forests)
        if not (n == len(engines) == len(seeds)) or n == 0:
            raise ValueError("one seed and one engine per forest, at least one forest")
        by_id: dict[int, tuple] = {}  # one parameter block per distinct engine object
        built = [by_id[id(e)] if id(e) in by_id else by_id.setdefault(id(e), A.engine_params(**e)) for e in engines]
        fs = (C.c_void_p * n)(*[f.h for f in forests])
        ps = (C.POINTER(A.EngineParams) * n)(*[C.pointer(b[0]) for b in built])
        sd = np.ascontiguousarray(seeds, np.uint64)
        hs = (C.c_void_p * n)()
        check(lib().psx_place_mutations_batch(ctx.h, genome.h, n, fs, ps, A.ptr(sd, C.c_uint64), hs), ctx.h)
        del built
        return [cls(ctx, C.c_void_p(hs[k]), int(engines[k].get("ploidy", 2))) for k in range(n)]

    @classmethod
    def upload(cls, ctx: Context, forest: Forest, recs: np.ndarray, ploidy: int = 2) -> "Mutations":
        r = np.ascontiguousarray(recs, A.MUTATION_DTYPE)
        h = C.c_void_p()
        check(lib().psx_mutations_upload(ctx.h, forest.h, r.ctypes.data_as(C.c_void_p), len(r), ploidy,
                                         C.byref(h)), ctx.h)
        return cls(ctx, h, ploidy)

    def save(self, path) -> None:
        """Writes the table (records, copy-number events, chromosome lengths,
        ploidy, leaves) to a file (psx_mutations_save)."""
        check(lib().psx_mutations_save(self.ctx.h, self.h, os.fsencode(path)), self.ctx.h)

    @classmethod
    def load(cls, ctx: Context, path) -> "Mutations":
        """A table written by save, back in device memory (psx_mutations_load)."""
        h = C.c_void_p()
        check(lib().psx_mutations_load(ctx.h, os.fsencode(path), C.byref(h)), ctx.h)
        with open(path, "rb") as f:  # the header's ploidy (bytes 12..15; the library has checked the file)
            ploidy = int(np.frombuffer(f.read(16)[12:16], np.uint32)[0])
        return cls(ctx, h, ploidy)

    def cna(self) -> np.ndarray:
        """Copy-number events (CNA_EVENT_DTYPE), in DFS order of their node."""
        n = C.c_uint32()
        check(lib().psx_mutations_cna(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, A.CNA_EVENT_DTYPE)
        if n.value:
            check(lib().psx_mutations_cna(self.h, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
        return out

    def set_cna(self, events: np.ndarray) -> None:
        ev = np.ascontiguousarray(events, A.CNA_EVENT_DTYPE)
        check(lib().psx_mutations_set_cna(self.ctx.h, self.h, ev.ctypes.data_as(C.c_void_p) if len(ev) else None,
                                          len(ev)), self.ctx.h)

    def download(self) -> np.ndarray:
        out = np.zeros(self.n, A.MUTATION_DTYPE)
        check(lib().psx_mutations_download(self.ctx.h, self.h, out.ctypes.data_as(C.c_void_p), self.n), self.ctx.h)
        return out


def num_out_samples(forest: Forest, **seq) -> int:
    p, keep = A.seq_params(**seq)
    return int(lib().psx_seq_num_samples(forest.h, C.byref(p)))


def simulate_seq_arrays(ctx: Context, genome: Genome, forest: Forest, muts: Mutations, seed: int, **seq):
    """Count-only sequencing into host arrays: (alt, depth, present, n_reads),
    each [n_out_samples, n_mutations]."""
    p, keep = A.seq_params(**seq)
    S = int(lib().psx_seq_num_samples(forest.h, C.byref(p)))
    M = muts.n
    alt = np.zeros(S * M, np.int32)
    dep = np.zeros(S * M, np.int32)
    pres = np.zeros(S * M, np.uint8)
    nr = C.c_uint64()
    check(lib().psx_simulate_seq(ctx.h, genome.h, forest.h, muts.h, C.byref(p), C.c_uint64(seed),
                                 A.ptr(alt, C.c_int32), A.ptr(dep, C.c_int32), A.ptr(pres, C.c_uint8), C.byref(nr)),
          ctx.h)
    del keep
    return alt.reshape(S, M), dep.reshape(S, M), pres.reshape(S, M), nr.value


def num_reads(genome: Genome, forest: Forest, **seq) -> int:
    p, keep = A.seq_params(**seq)
    n = C.c_uint64()
    check(lib().psx_seq_num_reads(genome.h, forest.h, C.byref(p), C.byref(n)))
    return n.value


def num_reads_table(ctx: Context, genome: Genome, forest: Forest, muts: Mutations, seed: int, **seq) -> int:
    """Reads a sequencing call of this table and seed simulates, copy-number
    events included (psx_seq_num_reads_table)."""
    p, keep = A.seq_params(**seq)
    n = C.c_uint64()
    check(lib().psx_seq_num_reads_table(ctx.h, genome.h, forest.h, muts.h, C.byref(p), C.c_uint64(seed),
                                        C.byref(n)), ctx.h)
    return n.value


class SeqJob:
    """One job of a batched launch, with device-resident outputs."""

    def __init__(self, ctx: Context, genome: Genome, forest: Forest, muts: Mutations, seed: int, **seq):
        self.genome, self.forest, self.muts, self.seed, self.seq = genome, forest, muts, int(seed), seq
        self.n_out = num_out_samples(forest, **seq)
        shape = (self.n_out, max(muts.n, 1))
        self.alt = DeviceArray(ctx, shape, np.int32)
        self.depth = DeviceArray(ctx, shape, np.int32)
        self.present = DeviceArray(ctx, shape, np.uint8)

    def results(self):
        M = self.muts.n
        return (self.alt.numpy()[:, :M], self.depth.numpy()[:, :M], self.present.numpy()[:, :M])


def _job_array(jobs):
    arr = (A.SeqJob * len(jobs))()
    keep = []
    for i, j in enumerate(jobs):
        p, k = A.seq_params(**j.seq)
        keep.append(k)
        arr[i] = A.SeqJob(j.genome.h, j.forest.h, j.muts.h, p, j.seed, j.alt.ptr, j.depth.ptr, j.present.ptr)
    return arr, keep


def simulate_seq_batch(ctx: Context, jobs: list, config: A.LaunchConfig | None = None) -> int:
    """All jobs in one persistent launch; returns the simulated reads."""
    arr, keep = _job_array(jobs)
    nr = C.c_uint64()
    check(lib().psx_simulate_seq_batch(ctx.h, arr, len(jobs), None if config is None else C.byref(config),
                                       C.byref(nr)), ctx.h)
    return nr.value


class SeqPlan:
    """A batched launch prepared once (psx_seq_plan): run() replays it."""

    def __init__(self, ctx: Context, jobs: list, config: A.LaunchConfig | None = None):
        self.ctx, self.jobs = ctx, list(jobs)
        self._arr, self._keep = _job_array(self.jobs)
        h = C.c_void_p()
        check(lib().psx_seq_plan_create(ctx.h, self._arr, len(self.jobs), None if config is None else
                                        C.byref(config), C.byref(h)), ctx.h)
        self.h = h

    def run(self) -> int:
        nr = C.c_uint64()
        check(lib().psx_seq_plan_run(self.h, C.byref(nr)), self.ctx.h)
        return nr.value

    def close(self):
        if getattr(self, "h", None):
            lib().psx_seq_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(ctx: Context, jobs: list) -> tuple[A.LaunchConfig, float]:
    arr, keep = _job_array(jobs)
    best = A.LaunchConfig()
    ms = C.c_double()
    check(lib().psx_autotune_seq(ctx.h, arr, len(jobs), C.byref(best), C.byref(ms)), ctx.h)
    return best, ms.value


def simulate_seq_device(ctx: Context, genome: Genome, forest: Forest, muts: Mutations, seed: int, alt: DeviceArray,
                        depth: DeviceArray, present: DeviceArray, **seq) -> int:
    p, keep = A.seq_params(**seq)
    nr = C.c_uint64()
    check(lib().psx_simulate_seq_device(ctx.h, genome.h, forest.h, muts.h, C.byref(p), C.c_uint64(seed), alt.ptr,
                                        depth.ptr, present.ptr, C.byref(nr)), ctx.h)
    return nr.value


def simulate_seq_device_noalloc(ctx: Context, genome: Genome, forest: Forest, muts: Mutations, seed: int,
                                alt: DeviceArray, depth: DeviceArray, present: DeviceArray, **seq) -> int:
    """simulate_seq_device with no device allocation (descriptors in the
    library's fixed per-GPU scratch array; calls on one GPU must not overlap)."""
    p, keep = A.seq_params(**seq)
    nr = C.c_uint64()
    check(lib().psx_simulate_seq_device_noalloc(ctx.h, genome.h, forest.h, muts.h, C.byref(p), C.c_uint64(seed),
                                                alt.ptr, depth.ptr, present.ptr, C.byref(nr)), ctx.h)
    return nr.value


def seq_vaf_f32(ctx: Context, muts: Mutations, alt: DeviceArray, depth: DeviceArray,
                present: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
    """The result table's VAF in fp32 on the device for every record of the
    [n_out_samples, n_mutations] device arrays of simulate_seq_device (SID
    heads; 0 elsewhere), into `out` when given."""
    S = alt.shape[0]
    if out is None:
        out = DeviceArray(ctx, (S, max(muts.n, 1)), np.float32)
    if out.dtype != np.float32 or out.nbytes < 4 * S * muts.n:
        raise ValueError(f"the VAF buffer must hold {S} x {muts.n} float32 values")
    check(lib().psx_seq_vaf_f32(ctx.h, muts.h, S, alt.ptr, depth.ptr, present.ptr, out.ptr), ctx.h)
    return out


def read_words(**seq) -> int:
    p, keep = A.seq_params(**seq)
    return int(lib().psx_read_words(C.byref(p)))


def _check_read_buffer(out: DeviceArray, rw: int) -> None:
    """The library writes rows of rw uint32 words: a caller's buffer must be
    a [rows, rw] uint32 array, or rows would be written past its end."""
    if out.dtype != np.uint32 or len(out.shape) != 2 or out.shape[1] != rw:
        raise ValueError(f"the read buffer must be a [reads, {rw}] uint32 array (got {out.shape} {out.dtype})")


def simulate_seq_reads(ctx: Context, genome: Genome, forest: Forest, muts: Mutations, seed: int,
                       out: DeviceArray | None = None, **seq):
    """Materialised reads into a device buffer of shape [n_reads, read_words]."""
    n = num_reads_table(ctx, genome, forest, muts, seed, **seq)
    rw = read_words(**seq)
    if out is None:
        out = DeviceArray(ctx, (max(n, 1), rw), np.uint32)
    _check_read_buffer(out, rw)
    p, keep = A.seq_params(**seq)
    got = C.c_uint64()
    check(lib().psx_simulate_seq_reads(ctx.h, genome.h, forest.h, muts.h, C.byref(p), C.c_uint64(seed), out.ptr,
                                       out.shape[0], C.byref(got)), ctx.h)
    return out, got.value


def simulate_seq_reads_batch(ctx: Context, jobs: list, out: DeviceArray | None = None):
    """Materialised reads of every job in one launch: (buffer [n_reads,
    read_words], job_first with len(jobs) + 1 entries, n_reads); job k's reads
    are rows job_first[k]:job_first[k + 1]."""
    n = sum(num_reads_table(ctx, j.genome, j.forest, j.muts, j.seed, **j.seq) for j in jobs)
    rw = read_words(**jobs[0].seq) if jobs else 1
    if out is None:
        out = DeviceArray(ctx, (max(n, 1), rw), np.uint32)
    _check_read_buffer(out, rw)
    arr, keep = _job_array(jobs)
    first = (C.c_uint64 * (len(jobs) + 1))()
    got = C.c_uint64()
    check(lib().psx_simulate_seq_reads_batch(ctx.h, arr, len(jobs), out.ptr, out.shape[0], first, C.byref(got)),
          ctx.h)
    return out, [int(x) for x in first], got.value


def seq_table(recs: np.ndarray, alt, depth, present, include_non_sequenced: bool = False):
    """psx_seq_table: (row record indices, occ, cov, vaf) with [sample, row] arrays."""
    r = np.ascontiguousarray(recs, A.MUTATION_DTYPE)
    S, M = alt.shape
    Mx = max(M, 1)
    a = np.ascontiguousarray(alt, np.int32)
    d = np.ascontiguousarray(depth, np.int32)
    p = np.ascontiguousarray(present, np.uint8)
    rows = np.zeros(Mx, np.uint64)
    occ = np.zeros(S * Mx, np.int32)
    cov = np.zeros(S * Mx, np.int32)
    vaf = np.zeros(S * Mx, np.float64)
    nr = C.c_uint64()
    check(lib().psx_seq_table(r.ctypes.data_as(C.c_void_p), M, S, A.ptr(a, C.c_int32), A.ptr(d, C.c_int32),
                              A.ptr(p, C.c_uint8), int(bool(include_non_sequenced)), C.byref(nr),
                              A.ptr(rows, C.c_uint64), A.ptr(occ, C.c_int32), A.ptr(cov, C.c_int32),
                              A.ptr(vaf, C.c_double)))
    n = nr.value
    return (rows[:n].copy(), occ.reshape(S, Mx)[:, :n].copy(), cov.reshape(S, Mx)[:, :n].copy(),
            vaf.reshape(S, Mx)[:, :n].copy())


def probe_valu(ctx: Context, mode: str, iters: int) -> tuple[float, int]:
    """The vector-ALU ceiling kernels (psx_probe_valu): mode "philox" or
    "int"; returns (HIP-event ms of one launch, lanes x iterations)."""
    ms, n = C.c_double(), C.c_uint64()
    m = {"philox": A.PSX_PROBE_PHILOX, "int": A.PSX_PROBE_INT}[mode]
    check(lib().psx_probe_valu(ctx.h, m, int(iters), C.byref(ms), C.byref(n)), ctx.h)
    return ms.value, n.value


def probe_store(ctx: Context, out: "DeviceArray", mode: str) -> tuple[float, int]:
    """The HBM write-ceiling kernels (psx_probe_store) over out's memory
    (overwritten): mode "fill", "chunks" or "chunks_shifted"; returns
    (HIP-event ms of one launch, bytes written)."""
    ms, n = C.c_double(), C.c_uint64()
    m = {"fill": A.PSX_PROBE_STORE_FILL, "chunks": A.PSX_PROBE_STORE_CHUNKS,
         "chunks_shifted": A.PSX_PROBE_STORE_CHUNKS_SHIFTED}[mode]
    check(lib().psx_probe_store(ctx.h, C.c_void_p(out.ptr), out.nbytes, m, C.byref(ms), C.byref(n)), ctx.h)
    return ms.value, n.value


def synth_forest(*, growth, death, intro_at, n_cells: int, n_samples: int, sample_size: int, seed: int) -> dict:
    """Synthetic birth-death samples forest (host generator in libpsx)."""
    g = np.ascontiguousarray(growth, np.float64)
    d = np.ascontiguousarray(death, np.float64)
    ia = np.ascontiguousarray(intro_at, np.uint64)
    sp = A.SynthParams(len(g), A.ptr(g, C.c_double), A.ptr(d, C.c_double), A.ptr(ia, C.c_uint64), int(n_cells),
                       int(n_samples), int(sample_size))
    h = C.c_void_p()
    check(lib().psx_synth_forest_run(C.byref(sp), C.c_uint64(seed), C.byref(h)))
    try:
        n = lib().psx_synth_forest_nodes(h)
        out = dict(cell_id=np.zeros(n, np.int64), parent=np.zeros(n, np.int64), species=np.zeros(n, np.uint32),
                   sample=np.zeros(n, np.int32), birth=np.zeros(n, np.float64))
        check(lib().psx_synth_forest_get(h, A.ptr(out["cell_id"], C.c_int64), A.ptr(out["parent"], C.c_int64),
                                         A.ptr(out["species"], C.c_uint32), A.ptr(out["sample"], C.c_int32),
                                         A.ptr(out["birth"], C.c_double)))
    finally:
        lib().psx_synth_forest_free(h)
    out["n_samples"] = int(n_samples)
    return out


class Server:
    """Serving runtime (psx_server): continuous batching of count-only
    sequencing requests within an HBM budget, results copied to host arrays."""

    def __init__(self, ctx: Context, hbm_budget: int, max_inflight: int = 2, max_batch: int = 16,
                 launch: A.LaunchConfig | None = None, overlap_after_us: int = 250):
        self.ctx = ctx
        if launch is None:
            # every lane launches a full persistent grid: a launch that starts
            # while another runs takes its compute units as they free up
            launch = A.LaunchConfig(256, 512, 1, 8, 0)
        cfg = A.ServerConfig(int(hbm_budget), int(max_inflight), int(max_batch), launch, int(overlap_after_us))
        h = C.c_void_p()
        check(lib().psx_server_create(ctx.h, C.byref(cfg), C.byref(h)), ctx.h)
        self.h = h
        self._live: dict[int, tuple] = {}

    def _err(self) -> str:
        m = lib().psx_server_last_error(self.h)
        return m.decode() if m else ""

    def submit(self, genome: Genome, forest: Forest, muts: Mutations, seed: int, **seq) -> int:
        p, keep = A.seq_params(**seq)
        S = int(lib().psx_seq_num_samples(forest.h, C.byref(p)))
        M = muts.n
        alt = np.zeros((S, M), np.int32)
        dep = np.zeros((S, M), np.int32)
        pres = np.zeros((S, M), np.uint8)
        r = A.Request(genome.h, forest.h, muts.h, p, int(seed), A.ptr(alt, C.c_int32), A.ptr(dep, C.c_int32),
                      A.ptr(pres, C.c_uint8))
        t = C.c_uint64()
        st = lib().psx_server_submit(self.h, C.byref(r), C.byref(t))
        if st != 0:
            raise PsxError(st, self._err())
        self._live[t.value] = (alt, dep, pres, keep)
        return t.value

    def wait(self, ticket: int, timeout_ms: float = -1.0):
        """(alt, depth, present, stats) of a finished request."""
        s = A.RequestStats()
        st = lib().psx_server_wait(self.h, C.c_uint64(ticket), float(timeout_ms), C.byref(s))
        if st == 4 and timeout_ms >= 0 and "timed out" in self._err():
            raise TimeoutError(self._err())  # the request stays queued
        alt, dep, pres, _ = self._live.pop(ticket)
        if st != 0:
            raise PsxError(st, self._err())
        stats = dict(n_reads=s.n_reads, batch_size=s.batch_size, lane=s.lane, queued_ms=s.queued_ms,
                     latency_ms=s.latency_ms)
        return alt, dep, pres, stats

    def totals(self) -> dict:
        t = A.ServerTotals()
        check(lib().psx_server_totals_get(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in A.ServerTotals._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().psx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
