// Internal declarations shared by the host and device translation units of
// libpsx. Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/psx.h"
#include "psx_rng.h"

// -------------------------------------------------------------- status/error

struct psx_error : std::exception {
    psx_status code;
    std::string msg;
    psx_error(psx_status c, std::string m) : code(c), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

[[noreturn]] void psx_throw(psx_status code, const char* fmt, ...);

#define PSX_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            psx_throw(e_ == hipErrorOutOfMemory ? PSX_ENOMEM_DEVICE : PSX_EHIP,          \
                      "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__,   \
                      __LINE__);                                                         \
    } while (0)

// Runs fn, converting exceptions into a status and the context message.
template <class F>
psx_status psx_guard(psx_ctx* ctx, F&& fn);

// -------------------------------------------------------------- device memory

struct dbuf {
    void* p = nullptr;
    size_t bytes = 0;
    dbuf() = default;
    dbuf(const dbuf&) = delete;
    dbuf& operator=(const dbuf&) = delete;
    dbuf(dbuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    dbuf& operator=(dbuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~dbuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) n = 256;
        PSX_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    // Grows to at least n bytes, keeping nothing.
    void reserve(size_t n) { if (n > bytes) alloc(n); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// ------------------------------------------------------------------- handles

struct psx_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    double last_kernel_ms = 0.0;
    double last_place_ms = 0.0;
    dbuf scratch;       // per-call tables and descriptors
    dbuf queue;         // persistent-kernel work counters
    dbuf place_nodes;   // placement temporaries per node and per mutation (grown, kept for later calls)
    dbuf place_muts;
    dbuf place_tab;     // the job table and round state of the last placement call
    dbuf seq_out;       // psx_simulate_seq's device tables before their copy to the caller (grown, kept)
    uint8_t* noalloc_scratch = nullptr;  // device address of the device's no-allocation scratch
    uint8_t* host_io = nullptr;          // pinned host memory for small device-to-host readbacks
    size_t host_io_bytes = 0;
    uint8_t* host_stage = nullptr;       // pinned chunks of large copies to pageable memory
    size_t host_stage_bytes = 0;
    bool tuned = false;
    psx_launch_config tuned_cfg{};
    std::atomic<bool> cancel{false};  // psx_ctx_cancel: the running call stops at its next work-chunk boundary
};

// Throws PSX_ECANCELLED when a cancellation was requested (and consumes the request).
inline void psx_check_cancel(psx_ctx* ctx) {
    if (ctx->cancel.exchange(false)) psx_throw(PSX_ECANCELLED, "cancelled");
}

// ctx's pinned readback area, at least bytes long (grown after a stream synchronisation, never shrunk).
uint8_t* psx_host_io(psx_ctx* ctx, size_t bytes);

// Device-to-host copies into pageable memory through pinned chunks: the
// ranges are cut into 2 MiB pieces dealt to up to 8 host threads, each moving
// its pieces one by one (a DMA on ctx's stream into its pinned chunk, then a
// memcpy out), so one thread's DMA overlaps the others' memcpys. Returns when
// every byte is in place; the stream must hold nothing the copies should not
// wait for.
struct psx_copy_range {
    const void* src;  // device
    void* dst;        // host
    size_t bytes;
};
void psx_copy_out(psx_ctx* ctx, const std::vector<psx_copy_range>& ranges);

// Device address of the no-allocation entry point's scratch array on
// `device`, which must be the current device; looked up once per device.
uint8_t* psx_noalloc_scratch_init(int device);

// Indel (ID83) classes of a genome's reference: sites per class and, for the
// classes too sparse for a rejection search, their sites listed (global
// position << 10 | payload, ascending per class). Built on first use.
struct psx_id_index {
    uint64_t count[PSX_ID_CLASSES] = {};
    uint32_t listed[PSX_ID_CLASSES] = {};
    uint64_t off[PSX_ID_CLASSES + 1] = {};
    dbuf d_list;
    double build_ms = 0.0;
};

// SBS context index of a genome's reference at one sampling s: per
// pyrimidine-centred trinucleotide context (32), the sites (positions q with
// 1 <= q < len - 1) counted, and every s-th one in genome order from the
// first kept as a 32-bit global position, ascending per context. d_meta holds
// kept[32] then off[33] (uint64) on the device. Built on first use.
struct psx_sbs_index {
    uint32_t sampling = 1;
    uint64_t count[32] = {};
    uint64_t kept[32] = {};
    uint64_t off[33] = {};
    dbuf d_list;
    dbuf d_meta;
    double build_ms = 0.0;
    // with genomes below 2^31 bases, bit 31 of each list entry holds the
    // site's strand (set: the reference base there is G or A, the purine), so
    // a draw needs no read of the reference
    bool strand_bit = false;
};

struct psx_genome {
    psx_ctx* ctx = nullptr;
    uint32_t n_chr = 0;
    std::vector<uint32_t> len;
    std::vector<uint64_t> woff;   // first word of each chromosome
    std::vector<uint64_t> boff;   // absolute base offset, n_chr + 1 entries
    uint64_t words = 0;
    dbuf packed;
    std::mutex id_mu;
    std::unique_ptr<psx_id_index> id_index;
    std::mutex sbs_mu;
    std::map<uint32_t, std::unique_ptr<psx_sbs_index>> sbs_index;  // by sampling
};

// The genome's indel index, built on ctx's stream at the first call.
const psx_id_index& psx_genome_id_index(psx_ctx* ctx, psx_genome* g);
// The genome's SBS context index at a sampling (0 counts as 1), built on ctx's
// stream at the first call with that sampling.
const psx_sbs_index& psx_genome_sbs_index(psx_ctx* ctx, psx_genome* g, uint32_t sampling);

struct psx_forest {
    psx_ctx* ctx = nullptr;
    uint32_t n_nodes = 0, n_leaves = 0, n_samples = 0;
    // host copies of the flattened forest (DFS pre-order)
    std::vector<int64_t> cell_id;
    std::vector<uint32_t> parent, species, leaf_lo, leaf_hi;
    uint32_t max_species = 0;
    std::vector<double> birth;
    std::vector<uint32_t> sample_begin, sample_leaves;
    // device copies
    dbuf d_leaf_lo, d_leaf_hi, d_species, d_parent, d_sample_leaves;
    // per sample, per block of 64 leaf ranks: {sample leaves below the block, the block's membership bits (low,
    // high word), 0}; leaf_blocks = n_leaves / 64 + 1 entries per sample, so rank n_leaves has an entry too
    uint32_t leaf_blocks = 0;
    dbuf d_leaf_count;
};

// Per-stratum index of a mutation table for one set of chromosome lengths.
struct psx_mut_index {
    std::vector<uint32_t> chr_len;
    std::vector<uint64_t> bucket_off;   // per chromosome offset into bucket
    std::shared_ptr<dbuf> store;        // holds bucket (the indexes of a batched placement share one block)
    uint32_t* bucket = nullptr;         // per chromosome: first record at or after k * PSX_STRATUM
};

// Bucket offsets of a stratum index for the given chromosome lengths.
std::vector<uint64_t> psx_index_layout(const std::vector<uint32_t>& chr_len);

// Mutation table in SID order: device SoA for the kernels, AoS for download.
struct psx_mutations {
    psx_ctx* ctx = nullptr;
    uint64_t n = 0;
    uint32_t ploidy = 2;
    uint32_t n_leaves = 0;
    std::vector<uint64_t> chr_begin;    // PSX_MAX_CHROMOSOMES + 1
    std::vector<uint32_t> chr_len;      // chromosome lengths at placement (uploads: largest position + 1)
    // device storage: the records (psx_mutation[n]) and the kernels' SoA (uint32[n] each), carved out of one
    // block that the tables of one batched placement share
    std::shared_ptr<dbuf> store;
    psx_mutation* recs = nullptr;
    uint32_t *pos = nullptr, *lo = nullptr, *hi = nullptr, *info = nullptr;
    uint64_t* d_chr_begin = nullptr;    // chr_begin on the device
    // every stratum index built so far, one per set of chromosome lengths;
    // kept until the table is freed, since launches in flight may read any
    std::mutex idx_mu;
    std::vector<std::unique_ptr<psx_mut_index>> index;
    std::vector<psx_cna_event> cna;     // copy-number events in DFS order of their node
};

// Device bytes of a table of n records (records and SoA, 256-byte aligned parts).
size_t psx_mutations_bytes(uint64_t n);
// Points m's records and SoA into block at byte offset off (256-aligned).
void psx_mutations_carve(psx_mutations* m, const std::shared_ptr<dbuf>& block, size_t off);

// Validates copy-number events against a forest and the chromosome lengths.
void psx_cna_check(const std::vector<psx_cna_event>& ev, uint32_t n_leaves, uint32_t ploidy,
                   const std::vector<uint32_t>& chr_len);

// info word of a device mutation record (built by k_mut_soa)
#define PSX_INFO_ALLELE(i) ((i) & 0xFFu)
#define PSX_INFO_REF(i) (((i) >> 8) & 3u)
#define PSX_INFO_ALT(i) (((i) >> 10) & 3u)
#define PSX_INFO_KIND(i) (((i) >> 12) & 3u)
#define PSX_INFO_NATURE(i) (((i) >> 14) & 3u)
#define PSX_INFO_HEAD(i) (((i) >> 16) & 1u)
#define PSX_INFO_SOLO(i) (((i) >> 17) & 1u)  // the only record at its position

// ------------------------------------------------------------ host helpers

// Integer inversion table of a pmf (psx_rng.h): T[k] = floor(cdf_k * 2^32),
// T[n-1] = 2^32.
void psx_cdf_table(const double* pmf, uint32_t n, uint64_t* T);
struct psx_itable {
    uint32_t k0 = 0;
    std::vector<uint64_t> T;
};
psx_itable psx_poisson_table(double lambda);
psx_itable psx_binomial_table(uint32_t t, double p, uint32_t cap);
// Sequencing errors are drawn per template over its N = rpt L bases: the
// rate is accepted when P(Binomial(N, e) > PSX_ERR_KCAP) < 2^-40; the count
// is Binomial(N, e) on [0, min(N, PSX_ERR_KCAP)].
bool psx_error_rate_ok(uint32_t N, double e);
psx_itable psx_error_table(uint32_t N, double e);
psx_itable psx_insert_table(uint32_t mean, uint32_t sd);
// Kernel form of an inversion table (psx_guided_pick in psx_rng.h): draws
// are identical to psx_cdf_pick on the full table.
struct psx_ctable {
    uint32_t k0 = 0;
    std::vector<uint32_t> T;
    std::vector<uint16_t> guide;  // PSX_GUIDE entries
};
psx_ctable psx_compact_table(const psx_itable& t);

uint64_t psx_sid_key(const psx_mutation& r);

// Sorts recs into SID order, marks sid_head and builds the device table.
void psx_mutations_build(psx_ctx* ctx, psx_mutations* m, std::vector<psx_mutation>& recs,
                         const std::vector<uint32_t>& chr_len);
// m->recs holds m->n records in SID order: builds the SoA arrays and the
// per-stratum index for the chromosome lengths in m->chr_len.
void psx_mutations_finish(psx_ctx* ctx, psx_mutations* m, const std::vector<uint64_t>& chr_begin);
// The same with m->d_chr_begin already written on ctx's stream; with sync = false the caller synchronises ctx's
// stream before the table is used (and sets the host chr_begin).
void psx_mutations_finish_device(psx_ctx* ctx, psx_mutations* m, bool sync);
// The per-stratum index of m for chr_len. The first request for a set of
// lengths builds it on ctx's stream and synchronises that stream, so launches
// on any stream may read it at once; it stays valid until m is freed.
const psx_mut_index& psx_mutations_index(psx_ctx* ctx, psx_mutations* m, const std::vector<uint32_t>& chr_len,
                                         bool sync = true);
// The index of m for chr_len if one was built, else nullptr (allocates nothing).
const psx_mut_index* psx_mutations_find_index(psx_mutations* m, const std::vector<uint32_t>& chr_len);

// Copies host bytes into the context scratch arena at a 256-byte aligned
// offset and returns the device address.
struct scratch_builder {
    std::vector<uint8_t> host;
    size_t put(const void* src, size_t bytes);
    template <class T> size_t put_vec(const std::vector<T>& v) {
        return put(v.data(), v.size() * sizeof(T));
    }
};
void* psx_scratch_upload(psx_ctx* ctx, const scratch_builder& b);

// Message of the last failed call made without a context (psx_last_error(NULL)).
extern thread_local std::string psx_tls_err;

template <class F>
psx_status psx_guard(psx_ctx* ctx, F&& fn) {
    std::string& err = ctx ? ctx->err : psx_tls_err;
    try {
        fn();
        return PSX_OK;
    } catch (const psx_error& e) {
        err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        err = "host allocation failed";
        return PSX_ENOMEM_HOST;
    } catch (const std::exception& e) {
        err = e.what();
        return PSX_EINVAL;
    }
}
