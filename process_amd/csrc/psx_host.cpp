// Host side of libpsx: context, errors, integer sampling tables, forest
// flattening, the synthetic forest generator and the result table.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <numeric>
#include <thread>

#include "psx_internal.h"

// ------------------------------------------------------------------ errors

void psx_throw(psx_status code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw psx_error(code, buf);
}

// ----------------------------------------------------------------- context

extern "C" int psx_abi_version(void) { return PSX_ABI_VERSION; }

extern "C" psx_status psx_ctx_create(int device, psx_ctx** out) {
    if (!out) return PSX_EINVAL;
    psx_ctx* ctx = new psx_ctx();
    psx_status st = psx_guard(ctx, [&] {
        int n = 0;
        PSX_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) psx_throw(PSX_EINVAL, "no HIP device %d", device);
        ctx->device = device;
        PSX_HIP(hipSetDevice(device));
        hipDeviceProp_t prop;
        PSX_HIP(hipGetDeviceProperties(&prop, device));
        ctx->n_cu = prop.multiProcessorCount;
        PSX_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        PSX_HIP(hipEventCreate(&ctx->ev0));
        PSX_HIP(hipEventCreate(&ctx->ev1));
        ctx->noalloc_scratch = psx_noalloc_scratch_init(device);
    });
    if (st != PSX_OK) {
        // keep the context so the caller can read the message
        *out = ctx;
        return st;
    }
    *out = ctx;
    return PSX_OK;
}

extern "C" psx_status psx_ctx_trim(psx_ctx* ctx) {
    if (!ctx) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        ctx->scratch.release();
        ctx->place_nodes.release();
        ctx->place_muts.release();
        ctx->place_tab.release();
        ctx->seq_out.release();
        if (ctx->host_io) PSX_HIP(hipHostFree(ctx->host_io));
        ctx->host_io = nullptr;
        ctx->host_io_bytes = 0;
        if (ctx->host_stage) PSX_HIP(hipHostFree(ctx->host_stage));
        ctx->host_stage = nullptr;
        ctx->host_stage_bytes = 0;
    });
}

extern "C" void psx_ctx_destroy(psx_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        ctx->scratch.release();
        ctx->queue.release();
        ctx->place_nodes.release();
        ctx->place_muts.release();
        if (ctx->host_io) (void)hipHostFree(ctx->host_io);
        if (ctx->host_stage) (void)hipHostFree(ctx->host_stage);
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

thread_local std::string psx_tls_err;

extern "C" const char* psx_last_error(const psx_ctx* ctx) {
    return ctx ? ctx->err.c_str() : psx_tls_err.c_str();
}

extern "C" psx_status psx_ctx_synchronize(psx_ctx* ctx) {
    if (!ctx) return PSX_EINVAL;
    return psx_guard(ctx, [&] { PSX_HIP(hipStreamSynchronize(ctx->stream)); });
}

extern "C" psx_status psx_ctx_last_kernel_ms(const psx_ctx* ctx, double* ms) {
    if (!ctx || !ms) return PSX_EINVAL;
    *ms = ctx->last_kernel_ms;
    return PSX_OK;
}

extern "C" psx_status psx_ctx_cancel(psx_ctx* ctx) {
    if (!ctx) return PSX_EINVAL;
    ctx->cancel.store(true);
    return PSX_OK;
}

extern "C" psx_status psx_ctx_last_place_ms(const psx_ctx* ctx, double* ms) {
    if (!ctx || !ms) return PSX_EINVAL;
    *ms = ctx->last_place_ms;
    return PSX_OK;
}

extern "C" psx_status psx_ctx_mem_info(psx_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!ctx || !free_bytes || !total_bytes) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        size_t f = 0, t = 0;
        PSX_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = f;
        *total_bytes = t;
    });
}

// ----------------------------------------------------------- device buffers

extern "C" psx_status psx_dev_alloc(psx_ctx* ctx, uint64_t bytes, void** out) {
    if (!ctx || !out) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        *out = nullptr;
        PSX_HIP(hipMalloc(out, bytes ? bytes : 1));
    });
}

extern "C" void psx_dev_free(psx_ctx* ctx, void* p) {
    if (!p) return;
    if (ctx) (void)hipSetDevice(ctx->device);
    (void)hipFree(p);
}

extern "C" psx_status psx_dev_upload(psx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        if (bytes) PSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" psx_status psx_dev_download(psx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        if (bytes >= (uint64_t(8) << 20)) {  // large: through pinned chunks on several threads
            PSX_HIP(hipStreamSynchronize(ctx->stream));
            psx_copy_out(ctx, {psx_copy_range{src, dst, (size_t)bytes}});
            return;
        }
        if (bytes) PSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" psx_status psx_dev_fill_u8(psx_ctx* ctx, void* dst, uint8_t value, uint64_t bytes) {
    if (!ctx || (bytes && !dst)) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        if (bytes) PSX_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" psx_status psx_dev_copy(psx_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (bytes && (!dst || !src))) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        if (bytes) PSX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

// ---------------------------------------------------------- scratch arena

size_t scratch_builder::put(const void* src, size_t bytes) {
    size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + bytes);
    if (bytes) std::memcpy(host.data() + off, src, bytes);
    return off;
}

uint8_t* psx_host_io(psx_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->host_io_bytes) return ctx->host_io;
    PSX_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->host_io) PSX_HIP(hipHostFree(ctx->host_io));
    ctx->host_io = nullptr;
    ctx->host_io_bytes = 0;
    const size_t n = std::max<size_t>(bytes, 64 << 10);
    PSX_HIP(hipHostMalloc((void**)&ctx->host_io, n, hipHostMallocDefault));
    ctx->host_io_bytes = n;
    return ctx->host_io;
}

void psx_copy_out(psx_ctx* ctx, const std::vector<psx_copy_range>& ranges) {
    constexpr size_t CHUNK = size_t(2) << 20;
    struct piece {
        const uint8_t* src;
        uint8_t* dst;
        size_t bytes;
    };
    std::vector<piece> pieces;
    for (const psx_copy_range& r : ranges)
        for (size_t o = 0; o < r.bytes; o += CHUNK)
            pieces.push_back({(const uint8_t*)r.src + o, (uint8_t*)r.dst + o, std::min(CHUNK, r.bytes - o)});
    if (pieces.empty()) return;
    const unsigned T = (unsigned)std::min<size_t>(8, pieces.size());
    if (ctx->host_stage_bytes < T * CHUNK) {
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->host_stage) PSX_HIP(hipHostFree(ctx->host_stage));
        ctx->host_stage = nullptr;
        ctx->host_stage_bytes = 0;
        PSX_HIP(hipHostMalloc((void**)&ctx->host_stage, 8 * CHUNK, hipHostMallocDefault));
        ctx->host_stage_bytes = 8 * CHUNK;
    }
    std::vector<hipEvent_t> ev(T, nullptr);
    std::vector<std::exception_ptr> err(T);
    auto work = [&](unsigned t) {
        try {
            PSX_HIP(hipSetDevice(ctx->device));
            uint8_t* buf = ctx->host_stage + t * CHUNK;
            for (size_t k = t; k < pieces.size(); k += T) {
                PSX_HIP(hipMemcpyAsync(buf, pieces[k].src, pieces[k].bytes, hipMemcpyDeviceToHost, ctx->stream));
                PSX_HIP(hipEventRecord(ev[t], ctx->stream));
                PSX_HIP(hipEventSynchronize(ev[t]));
                std::memcpy(pieces[k].dst, buf, pieces[k].bytes);
            }
        } catch (...) {
            err[t] = std::current_exception();
        }
    };
    std::vector<std::thread> th;
    try {
        for (unsigned t = 0; t < T; ++t) PSX_HIP(hipEventCreateWithFlags(&ev[t], hipEventDisableTiming));
        for (unsigned t = 1; t < T; ++t) th.emplace_back(work, t);
        work(0);
        for (std::thread& x : th) x.join();
    } catch (...) {  // a thread that could not start: the started ones finish first
        for (std::thread& x : th)
            if (x.joinable()) x.join();
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        throw;
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    for (const std::exception_ptr& e : err)
        if (e) std::rethrow_exception(e);
}

void* psx_scratch_upload(psx_ctx* ctx, const scratch_builder& b) {
    ctx->scratch.reserve(b.host.size() + 256);
    if (!b.host.empty())
        PSX_HIP(hipMemcpyAsync(ctx->scratch.p, b.host.data(), b.host.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    return ctx->scratch.p;
}

// ------------------------------------------------------- sampling tables

void psx_cdf_table(const double* pmf, uint32_t n, uint64_t* T) {
    if (n == 0) psx_throw(PSX_EINVAL, "empty distribution");
    double total = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
        if (!(pmf[k] >= 0.0)) psx_throw(PSX_EINVAL, "negative or NaN probability");
        total += pmf[k];
    }
    if (!(total > 0.0) || !std::isfinite(total)) psx_throw(PSX_EINVAL, "distribution has no mass");
    double cum = 0.0;
    for (uint32_t k = 0; k + 1 < n; ++k) {
        cum += pmf[k];
        const double x = cum / total;
        T[k] = x >= 1.0 ? (uint64_t(1) << 32) : (uint64_t)(x * 4294967296.0);
    }
    T[n - 1] = uint64_t(1) << 32;
}

// Poisson(lambda) on [max(0, floor(l - 40 sqrt(l) - 10)), ceil(l + 40 sqrt(l) + 10)],
// log pmf = -l + k log l - lgamma(k + 1).
psx_itable psx_poisson_table(double lambda) {
    if (!(lambda >= 0.0) || !std::isfinite(lambda)) psx_throw(PSX_EINVAL, "invalid Poisson mean %g", lambda);
    psx_itable t;
    if (lambda == 0.0) {
        t.T.assign(1, uint64_t(1) << 32);
        return t;
    }
    const double sd = std::sqrt(lambda);
    double lo = std::floor(lambda - 40.0 * sd - 10.0);
    if (lo < 0.0) lo = 0.0;
    const double hi = std::ceil(lambda + 40.0 * sd + 10.0);
    const double cnt = hi - lo + 1.0;
    if (cnt > 65536.0) psx_throw(PSX_EINVAL, "Poisson mean %g too large", lambda);
    const uint32_t m = (uint32_t)cnt;
    std::vector<double> pmf(m);
    const double ll = std::log(lambda);
    for (uint32_t i = 0; i < m; ++i) {
        const double k = lo + (double)i;
        pmf[i] = std::exp(-lambda + k * ll - std::lgamma(k + 1.0));
    }
    t.k0 = (uint32_t)lo;
    t.T.resize(m);
    psx_cdf_table(pmf.data(), m, t.T.data());
    return t;
}

static psx_itable binomial_range(uint32_t n, double p, double lo, double hi, uint32_t cap) {
    const double cnt = hi - lo + 1.0;
    if (cnt > (double)cap) psx_throw(PSX_EINVAL, "binomial table too large");
    const uint32_t m = (uint32_t)cnt;
    std::vector<double> pmf(m);
    const double lp = std::log(p), lq = std::log1p(-p), lgt = std::lgamma((double)n + 1.0);
    for (uint32_t i = 0; i < m; ++i) {
        const double k = lo + (double)i;
        pmf[i] = std::exp(lgt - std::lgamma(k + 1.0) - std::lgamma((double)n - k + 1.0) + k * lp +
                          ((double)n - k) * lq);
    }
    psx_itable t;
    t.k0 = (uint32_t)lo;
    t.T.resize(m);
    psx_cdf_table(pmf.data(), m, t.T.data());
    return t;
}

psx_itable psx_binomial_table(uint32_t n, double p, uint32_t cap) {
    if (!(p > 0.0) || p > 1.0) psx_throw(PSX_EINVAL, "invalid binomial probability %g", p);
    if (p == 1.0) {
        psx_itable t;
        t.k0 = n;
        t.T.assign(1, uint64_t(1) << 32);
        return t;
    }
    const double m = (double)n * p, sd = std::sqrt((double)n * p * (1.0 - p));
    double lo = std::floor(m - 40.0 * sd - 10.0), hi = std::ceil(m + 40.0 * sd + 10.0);
    if (lo < 0.0) lo = 0.0;
    if (hi > (double)n) hi = (double)n;
    return binomial_range(n, p, lo, hi, cap);
}

bool psx_error_rate_ok(uint32_t N, double e) {
    if (N <= PSX_ERR_KCAP) return true;
    const double lp = std::log(e), lq = std::log1p(-e), lgn = std::lgamma((double)N + 1.0);
    double tail = 0.0;
    for (uint32_t k = PSX_ERR_KCAP + 1; k <= N; ++k) {
        const double term = std::exp(lgn - std::lgamma((double)k + 1.0) - std::lgamma((double)(N - k) + 1.0) +
                                     k * lp + (double)(N - k) * lq);
        tail += term;
        if (tail >= 0x1p-40 || (term < 1e-300 && (double)k > (double)N * e)) break;
    }
    return tail < 0x1p-40;
}

psx_itable psx_error_table(uint32_t N, double e) {
    const double hi = N < PSX_ERR_KCAP ? (double)N : (double)PSX_ERR_KCAP;
    return binomial_range(N, e, 0.0, hi, PSX_ERR_KCAP + 1);
}

// get_bin_dist (src/seq_simulation.cpp:431-451).
psx_itable psx_insert_table(uint32_t mean, uint32_t sd) {
    const double q = (double)(sd * sd) / (double)mean;
    const double p = 1.0 - q;
    if (!(p > 0.0))
        psx_throw(PSX_EINVAL,
                  "The insert size mean (%u) must be greater than or equal to its variance "
                  "(%u*%u=%u).\nSet the standard deviation and the variance by using the "
                  "optional parameter \"insert_size_stddev\".",
                  mean, sd, sd, sd * sd);
    const uint32_t t = (uint32_t)((double)mean / p);
    return psx_binomial_table(t, p, 4096);
}

psx_ctable psx_compact_table(const psx_itable& t) {
    const size_t n = t.T.size();
    const uint64_t one = uint64_t(1) << 32;
    size_t lo = 0;
    while (lo + 1 < n && t.T[lo] == 0) ++lo;  // thresholds of 0 are never reached
    size_t hi = lo;
    while (hi + 1 < n && t.T[hi] < one) ++hi;  // nothing after the first 2^32 is reached
    psx_ctable c;
    c.k0 = t.k0 + (uint32_t)lo;
    const size_t m = hi - lo + 1;
    c.T.resize(m);
    for (size_t i = 0; i < m; ++i) c.T[i] = t.T[lo + i] >= one ? 0xFFFFFFFFu : (uint32_t)t.T[lo + i];
    c.guide.resize(PSX_GUIDE);
    size_t k = 0;
    for (uint32_t g = 0; g < PSX_GUIDE; ++g) {
        const uint64_t target = (uint64_t)g << (32u - PSX_GUIDE_BITS);
        while (k + 1 < m && (uint64_t)c.T[k] <= target) ++k;
        c.guide[g] = (uint16_t)k;
    }
    return c;
}

extern "C" psx_status psx_insert_size_check(psx_ctx* ctx, uint32_t mean, uint32_t stddev) {
    return psx_guard(ctx, [&] {
        if (mean == 0) return;
        (void)psx_insert_table(mean, stddev);
    });
}

// --------------------------------------------------------------- SID order

uint64_t psx_sid_key(const psx_mutation& r) {
    const uint64_t payload = r.kind == PSX_SNV ? r.alt : (r.kind == PSX_INS ? r.seq : 0u);
    return (uint64_t(r.chr) << 58) | (uint64_t(r.pos) << 26) | (uint64_t(r.kind) << 24) |
           (uint64_t(r.len & 15u) << 20) | (payload & 0xFFFFFu);
}

// ------------------------------------------------------------------ forest

// SamplesForest node table (src/forest.hpp:29-67) flattened in DFS pre-order:
// roots and children by ascending cell id; a node's subtree holds the
// leaves with ranks [leaf_lo, leaf_hi).
static void flatten_forest(const psx_forest_desc* d, psx_forest* f) {
    const uint32_t n = d->n_nodes;
    if (n == 0) psx_throw(PSX_EINVAL, "the forest has no nodes");
    std::vector<uint32_t> by_id(n);
    std::iota(by_id.begin(), by_id.end(), 0u);
    std::sort(by_id.begin(), by_id.end(),
              [&](uint32_t a, uint32_t b) { return d->cell_id[a] < d->cell_id[b]; });
    for (uint32_t k = 1; k < n; ++k)
        if (d->cell_id[by_id[k]] == d->cell_id[by_id[k - 1]])
            psx_throw(PSX_EINVAL, "duplicated cell id %lld", (long long)d->cell_id[by_id[k]]);
    auto row_of = [&](int64_t id) -> int64_t {
        auto it = std::lower_bound(by_id.begin(), by_id.end(), id,
                                   [&](uint32_t r, int64_t v) { return d->cell_id[r] < v; });
        return (it != by_id.end() && d->cell_id[*it] == id) ? (int64_t)*it : -1;
    };
    std::vector<uint32_t> prow(n), first(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        if (d->parent[i] < 0) { prow[i] = PSX_NO_NODE; continue; }
        const int64_t r = row_of(d->parent[i]);
        if (r < 0) psx_throw(PSX_EINVAL, "unknown ancestor %lld", (long long)d->parent[i]);
        prow[i] = (uint32_t)r;
        first[r + 1]++;
    }
    std::partial_sum(first.begin(), first.end(), first.begin());
    std::vector<uint32_t> kids(n), fill(first.begin(), first.end() - 1);
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = by_id[k];
        if (prow[i] != PSX_NO_NODE) kids[fill[prow[i]]++] = i;
    }
    std::vector<uint32_t> order;  // rows in DFS pre-order
    order.reserve(n);
    std::vector<uint32_t> stack;
    for (uint32_t k = n; k-- > 0;)
        if (prow[by_id[k]] == PSX_NO_NODE) stack.push_back(by_id[k]);
    while (!stack.empty()) {
        const uint32_t i = stack.back();
        stack.pop_back();
        order.push_back(i);
        for (uint32_t c = first[i + 1]; c > first[i]; --c) stack.push_back(kids[c - 1]);
    }
    if (order.size() != n) psx_throw(PSX_EINVAL, "the ancestor relation has a cycle");
    std::vector<uint32_t> dfs(n);
    for (uint32_t v = 0; v < n; ++v) dfs[order[v]] = v;
    std::vector<uint32_t> size(n, 1);
    for (uint32_t v = n; v-- > 0;) {
        const uint32_t i = order[v];
        if (prow[i] != PSX_NO_NODE) size[dfs[prow[i]]] += size[v];
    }
    std::vector<uint32_t> leaves_before(n + 1, 0);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t i = order[v];
        leaves_before[v + 1] = leaves_before[v] + (first[i + 1] == first[i] ? 1u : 0u);
    }
    f->n_nodes = n;
    f->n_leaves = leaves_before[n];
    f->n_samples = d->n_samples;
    f->cell_id.resize(n); f->parent.resize(n); f->species.resize(n); f->birth.resize(n);
    f->leaf_lo.resize(n); f->leaf_hi.resize(n);
    f->sample_begin.assign(d->n_samples + 1, 0);
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t i = order[v];
        f->cell_id[v] = d->cell_id[i];
        f->parent[v] = prow[i] == PSX_NO_NODE ? PSX_NO_NODE : dfs[prow[i]];
        f->species[v] = d->species[i];
        f->max_species = std::max(f->max_species, d->species[i]);
        f->birth[v] = d->birth_time[i];
        f->leaf_lo[v] = leaves_before[v];
        f->leaf_hi[v] = leaves_before[v + size[v]];
        const int32_t s = d->sample[i];
        const bool leaf = first[i + 1] == first[i];
        if (s >= 0 && (!leaf || (uint32_t)s >= d->n_samples))
            psx_throw(PSX_EINVAL, "cell %lld: invalid sample %d", (long long)d->cell_id[i], s);
        if (s >= 0) f->sample_begin[s + 1]++;
    }
    std::partial_sum(f->sample_begin.begin(), f->sample_begin.end(), f->sample_begin.begin());
    f->sample_leaves.assign(f->n_leaves, 0);
    std::vector<uint32_t> sfill(f->sample_begin.begin(), f->sample_begin.end() - 1);
    for (uint32_t v = 0; v < n; ++v) {
        const int32_t s = d->sample[order[v]];
        if (s >= 0) f->sample_leaves[sfill[s]++] = leaves_before[v];
    }
    f->sample_leaves.resize(f->sample_begin[d->n_samples]);
}

template <class T>
static void upload_vec(psx_ctx* ctx, dbuf& b, const std::vector<T>& v) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty())
        PSX_HIP(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
}

extern "C" psx_status psx_forest_upload(psx_ctx* ctx, const psx_forest_desc* desc, psx_forest** out) {
    if (!ctx || !desc || !out) return PSX_EINVAL;
    psx_forest* f = new psx_forest();
    f->ctx = ctx;
    psx_status st = psx_guard(ctx, [&] {
        flatten_forest(desc, f);
        PSX_HIP(hipSetDevice(ctx->device));
        upload_vec(ctx, f->d_leaf_lo, f->leaf_lo);
        upload_vec(ctx, f->d_leaf_hi, f->leaf_hi);
        upload_vec(ctx, f->d_species, f->species);
        upload_vec(ctx, f->d_parent, f->parent);
        upload_vec(ctx, f->d_sample_leaves, f->sample_leaves);
        // membership of every leaf rank in every sample, with running counts: whether a sample holds a leaf of
        // a rank range [lo, hi) is count(hi) > count(lo), two lookups
        f->leaf_blocks = f->n_leaves / 64u + 1u;
        std::vector<uint32_t> lc((size_t)f->n_samples * f->leaf_blocks * 4u, 0u);
        for (uint32_t smp = 0; smp < f->n_samples; ++smp) {
            uint32_t* e = lc.data() + (size_t)smp * f->leaf_blocks * 4u;
            for (uint32_t k = f->sample_begin[smp]; k < f->sample_begin[smp + 1]; ++k) {
                const uint32_t r = f->sample_leaves[k];
                e[(r >> 6) * 4u + 1u + ((r >> 5) & 1u)] |= 1u << (r & 31u);
            }
            uint32_t run = 0;
            for (uint32_t b = 0; b < f->leaf_blocks; ++b) {
                e[b * 4u] = run;
                run += (uint32_t)__builtin_popcount(e[b * 4u + 1u]) + (uint32_t)__builtin_popcount(e[b * 4u + 2u]);
            }
        }
        upload_vec(ctx, f->d_leaf_count, lc);
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (st != PSX_OK) { delete f; return st; }
    *out = f;
    return PSX_OK;
}

extern "C" psx_status psx_forest_info(const psx_forest* f, uint32_t* n_nodes, uint32_t* n_leaves,
                                      uint32_t* n_samples) {
    if (!f) return PSX_EINVAL;
    if (n_nodes) *n_nodes = f->n_nodes;
    if (n_leaves) *n_leaves = f->n_leaves;
    if (n_samples) *n_samples = f->n_samples;
    return PSX_OK;
}

extern "C" psx_status psx_forest_flat(const psx_forest* f, int64_t* dfs_cell_id, uint32_t* parent_dfs,
                                      uint32_t* leaf_lo, uint32_t* leaf_hi, uint32_t* sample_begin,
                                      uint32_t* sample_leaves) {
    if (!f) return PSX_EINVAL;
    auto cp = [](auto* dst, const auto& v) { if (dst) std::copy(v.begin(), v.end(), dst); };
    cp(dfs_cell_id, f->cell_id);
    cp(parent_dfs, f->parent);
    cp(leaf_lo, f->leaf_lo);
    cp(leaf_hi, f->leaf_hi);
    cp(sample_begin, f->sample_begin);
    cp(sample_leaves, f->sample_leaves);
    return PSX_OK;
}

extern "C" void psx_forest_free(psx_forest* f) {
    if (!f) return;
    if (f->ctx) (void)hipSetDevice(f->ctx->device);
    delete f;
}

// ------------------------------------------------- synthetic birth-death forest

struct psx_synth_forest {
    std::vector<int64_t> cell_id, parent;
    std::vector<uint32_t> species;
    std::vector<int32_t> sample;
    std::vector<double> birth;
};

namespace {
struct draw_stream {
    uint64_t key;
    uint32_t attempt, purpose;
    psx_u4 at(uint64_t i, uint32_t sub = 0) const {
        return psx_draw(key, (uint32_t)i, (uint32_t)(i >> 32), attempt, (purpose << 24) | sub);
    }
};
double unit53(uint32_t hi, uint32_t lo) {
    // (0, 1): 53 random bits plus half an ulp
    const uint64_t b = ((uint64_t(hi) << 32) | lo) >> 11;
    return ((double)b + 0.5) * (1.0 / 9007199254740992.0);
}
}  // namespace

// Gillespie birth-death process with species introduction, then uniform
// disjoint samples and pruning to the ancestors of the sampled cells. Stands
// in for SpatialSimulation + get_samples_forest (src/simulation.cpp:1480-1483).
static void synth_run(const psx_synth_params* p, uint64_t seed, psx_synth_forest* out) {
    const uint32_t S = p->n_species;
    if (S == 0 || S > 64) psx_throw(PSX_EINVAL, "invalid number of species");
    if (p->n_cells < 1) psx_throw(PSX_EINVAL, "n_cells must be positive");
    if ((uint64_t)p->n_samples * p->sample_size > p->n_cells)
        psx_throw(PSX_EINVAL, "samples exceed the grown population");
    for (uint32_t s = 0; s < S; ++s)
        if (!(p->growth[s] > 0.0) || !(p->death[s] >= 0.0)) psx_throw(PSX_EINVAL, "invalid rates");
    for (uint32_t attempt = 0; attempt < 1000; ++attempt) {
        std::vector<int64_t> parent;
        std::vector<uint32_t> species;
        std::vector<double> birth;
        std::vector<std::vector<uint32_t>> alive(S);
        std::vector<uint32_t> slot;  // position of a live cell in its species list
        parent.push_back(-1); species.push_back(0); birth.push_back(0.0); slot.push_back(0);
        alive[0].push_back(0);
        std::vector<bool> introduced(S, false);
        introduced[0] = true;
        uint64_t n_alive = 1, event = 0, n_intro = 0;
        double t = 0.0;
        const draw_stream ev{seed, attempt, PSX_P_FOREST};
        auto remove_alive = [&](uint32_t cell) {
            auto& a = alive[species[cell]];
            const uint32_t k = slot[cell];
            a[k] = a.back();
            slot[a[k]] = k;
            a.pop_back();
        };
        auto add_cell = [&](int64_t par, uint32_t s) {
            const uint32_t id = (uint32_t)parent.size();
            parent.push_back(par); species.push_back(s); birth.push_back(t);
            slot.push_back((uint32_t)alive[s].size());
            alive[s].push_back(id);
        };
        bool extinct = false;
        while (n_alive < p->n_cells) {
            for (uint32_t s = 1; s < S; ++s) {
                if (introduced[s] || n_alive < p->intro_at[s] || alive[0].empty()) continue;
                const psx_u4 w = ev.at(n_intro++, 1);
                const uint32_t cell = alive[0][psx_uniform32(w.x, (uint32_t)alive[0].size())];
                remove_alive(cell);
                add_cell(cell, s);
                introduced[s] = true;
            }
            double R = 0.0;
            for (uint32_t s = 0; s < S; ++s) R += (double)alive[s].size() * (p->growth[s] + p->death[s]);
            if (R <= 0.0) { extinct = true; break; }
            const psx_u4 w = ev.at(event++);
            t += -std::log(unit53(w.x, w.y)) / R;
            double u = (double)w.z * (1.0 / 4294967296.0) * R;
            uint32_t s = 0;
            bool divide = true;
            for (; s < S; ++s) {
                const double nb = (double)alive[s].size() * p->growth[s];
                const double nd = (double)alive[s].size() * p->death[s];
                if (u < nb) { divide = true; break; }
                u -= nb;
                if (u < nd) { divide = false; break; }
                u -= nd;
            }
            if (s == S) {  // rounding at the top of the range
                s = S - 1;
                while (alive[s].empty()) --s;
                divide = true;
            }
            const uint32_t cell = alive[s][psx_uniform32(w.w, (uint32_t)alive[s].size())];
            remove_alive(cell);
            if (divide) {
                add_cell(cell, s);
                add_cell(cell, s);
                n_alive += 1;
            } else {
                n_alive -= 1;
                if (n_alive == 0) { extinct = true; break; }
            }
        }
        if (extinct) continue;
        // uniform disjoint samples by partial Fisher-Yates over the live cells
        std::vector<uint32_t> pool;
        pool.reserve(n_alive);
        for (uint32_t s = 0; s < S; ++s) pool.insert(pool.end(), alive[s].begin(), alive[s].end());
        std::sort(pool.begin(), pool.end());
        const draw_stream pk{seed, attempt, PSX_P_PICK};
        const uint64_t need = (uint64_t)p->n_samples * p->sample_size;
        std::vector<int32_t> sample_of(parent.size(), -1);
        for (uint64_t k = 0; k < need; ++k) {
            const psx_u4 w = pk.at(k);
            const uint64_t j = k + psx_uniform64(w.x, w.y, pool.size() - k);
            std::swap(pool[k], pool[j]);
            sample_of[pool[k]] = (int32_t)(k / p->sample_size);
        }
        std::vector<bool> keep(parent.size(), false);
        for (uint64_t k = 0; k < need; ++k) {
            int64_t c = pool[k];
            while (c >= 0 && !keep[c]) { keep[c] = true; c = parent[c]; }
        }
        for (size_t i = 0; i < parent.size(); ++i) {
            if (!keep[i]) continue;
            out->cell_id.push_back((int64_t)i);
            out->parent.push_back(parent[i]);
            out->species.push_back(species[i]);
            out->sample.push_back(sample_of[i]);
            out->birth.push_back(birth[i]);
        }
        return;
    }
    psx_throw(PSX_EINVAL, "the population went extinct in 1000 attempts");
}

extern "C" psx_status psx_synth_forest_run(const psx_synth_params* p, uint64_t seed, psx_synth_forest** out) {
    if (!p || !out) return PSX_EINVAL;
    psx_synth_forest* s = new psx_synth_forest();
    psx_status st = psx_guard(nullptr, [&] { synth_run(p, seed, s); });
    if (st != PSX_OK) { delete s; return st; }
    *out = s;
    return PSX_OK;
}

extern "C" uint32_t psx_synth_forest_nodes(const psx_synth_forest* s) {
    return s ? (uint32_t)s->cell_id.size() : 0u;
}

extern "C" psx_status psx_synth_forest_get(const psx_synth_forest* s, int64_t* cell_id, int64_t* parent,
                                           uint32_t* species, int32_t* sample, double* birth_time) {
    if (!s) return PSX_EINVAL;
    std::copy(s->cell_id.begin(), s->cell_id.end(), cell_id);
    std::copy(s->parent.begin(), s->parent.end(), parent);
    std::copy(s->species.begin(), s->species.end(), species);
    std::copy(s->sample.begin(), s->sample.end(), sample);
    std::copy(s->birth.begin(), s->birth.end(), birth_time);
    return PSX_OK;
}

extern "C" void psx_synth_forest_free(psx_synth_forest* s) { delete s; }

// ------------------------------------------------------------ result table

// get_result_dataframe (src/seq_simulation.cpp:169-181): rows are the SIDs
// active in some sample (get_active_mutations, :142-167); per sample the
// occurrences, the coverage at the SID position and VAF = occ/cov when the
// sample holds the SID, else 0 (add_sample_statistics, :92-140).
extern "C" psx_status psx_seq_table(const psx_mutation* recs, uint64_t n, uint32_t S,
                                    const int32_t* alt, const int32_t* depth, const uint8_t* present,
                                    int32_t include_non_sequenced, uint64_t* n_rows, uint64_t* row_rec,
                                    int32_t* occ, int32_t* cov, double* vaf) {
    if (!recs && n) return PSX_EINVAL;
    if (!n_rows) return PSX_EINVAL;
    std::vector<uint8_t> held(S);
    uint64_t rows = 0;
    uint64_t i = 0;
    while (i < n) {
        uint64_t end = i + 1;
        while (end < n && !recs[end].sid_head) ++end;
        bool active = false, any_held = false;
        for (uint32_t s = 0; s < S; ++s) {
            const uint64_t base = (uint64_t)s * n;
            bool h = alt[base + i] > 0;
            active |= h;
            for (uint64_t q = i; q < end && !h; ++q) h = present[base + q] != 0;
            held[s] = h;
            any_held |= h;
        }
        if (active || (include_non_sequenced && any_held)) {
            row_rec[rows] = i;
            for (uint32_t s = 0; s < S; ++s) {
                const uint64_t base = (uint64_t)s * n;
                const int32_t o = alt[base + i], c = depth[base + i];
                occ[base + rows] = held[s] ? o : 0;
                cov[base + rows] = c;
                vaf[base + rows] = held[s] ? (double)o / (double)c : 0.0;
            }
            ++rows;
        }
        i = end;
    }
    *n_rows = rows;
    return PSX_OK;
}
