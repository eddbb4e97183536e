// Device-resident reference genome (2-bit packed) and the device layout of
// mutation tables (SoA arrays plus a per-stratum index).
#include <algorithm>
#include <numeric>

#include "psx_internal.h"

namespace {

struct genome_synth_args {
    uint32_t* packed;
    uint64_t seed;
    uint32_t n_chr;
    uint32_t len[PSX_MAX_CHROMOSOMES];
    uint64_t woff[PSX_MAX_CHROMOSOMES];
    uint64_t qoff[PSX_MAX_CHROMOSOMES + 1];  // first Philox block of each chromosome
};

// One thread per Philox block: 4 packed words = 64 bases. Word w of
// chromosome c is word (w % 4) of Philox(seed, (w / 4, c, 0, REF << 24)).
__global__ void __launch_bounds__(256) k_genome_synth(const genome_synth_args a) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.qoff[a.n_chr]) return;
    uint32_t c = 0;
    while (q >= a.qoff[c + 1]) ++c;
    const uint64_t lb = q - a.qoff[c];
    const psx_u4 w = psx_draw(a.seed, (uint32_t)lb, c, 0u, PSX_P_REF << 24);
    const uint64_t words = ((uint64_t)a.len[c] + 15u) / 16u;
    const uint32_t rem = a.len[c] % 16u;
    for (int i = 0; i < 4; ++i) {
        const uint64_t wi = lb * 4u + (uint64_t)i;
        if (wi >= words) break;
        uint32_t v = psx_word(w, i);
        if (wi + 1 == words && rem) v &= (1u << (2u * rem)) - 1u;
        a.packed[a.woff[c] + wi] = v;
    }
}

__global__ void __launch_bounds__(256) k_mut_soa(const psx_mutation* __restrict__ recs, uint64_t n,
                                                 uint32_t* __restrict__ pos, uint32_t* __restrict__ lo,
                                                 uint32_t* __restrict__ hi, uint32_t* __restrict__ info) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const psx_mutation r = recs[i];
    pos[i] = r.pos;
    lo[i] = r.leaf_lo;
    hi[i] = r.leaf_hi;
    // alone at its position (records are in SID order: by chromosome, then position)
    const bool solo = (i == 0 || recs[i - 1].chr != r.chr || recs[i - 1].pos != r.pos) &&
                      (i + 1 == n || recs[i + 1].chr != r.chr || recs[i + 1].pos != r.pos);
    info[i] = uint32_t(r.allele) | (uint32_t(r.ref & 3u) << 8) | (uint32_t(r.alt & 3u) << 10) |
              (uint32_t(r.kind & 3u) << 12) | (uint32_t(r.nature & 3u) << 14) |
              (uint32_t(r.sid_head & 1u) << 16) | (uint32_t(solo) << 17);
}

struct bucket_args {
    const uint32_t* pos;
    uint32_t* bucket;
    uint32_t n_chr;
    const uint64_t* chr_begin;                 // [PSX_MAX_CHROMOSOMES + 1] on the device
    uint64_t boff[PSX_MAX_CHROMOSOMES + 1];  // first bucket entry of each chromosome
};

// bucket[boff[c] + k] = first record of chromosome c with pos >= k * PSX_STRATUM
// (as an absolute record index).
__global__ void __launch_bounds__(256) k_mut_bucket(const bucket_args a) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= a.boff[a.n_chr]) return;
    uint32_t c = 0;
    while (q >= a.boff[c + 1]) ++c;
    const uint64_t k = q - a.boff[c];
    const uint64_t target = k * PSX_STRATUM;
    uint64_t lo = a.chr_begin[c], hi = a.chr_begin[c + 1];
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)a.pos[mid] < target) lo = mid + 1; else hi = mid;
    }
    a.bucket[q] = (uint32_t)lo;
}

}  // namespace

static void genome_layout(psx_genome* g, uint32_t n_chr, const uint32_t* chr_len) {
    if (n_chr == 0 || n_chr > PSX_MAX_CHROMOSOMES)
        psx_throw(PSX_EINVAL, "the genome must have 1 to %u chromosomes", PSX_MAX_CHROMOSOMES);
    g->n_chr = n_chr;
    g->len.assign(chr_len, chr_len + n_chr);
    g->woff.assign(n_chr, 0);
    g->boff.assign(n_chr + 1, 0);
    uint64_t w = 0;
    for (uint32_t c = 0; c < n_chr; ++c) {
        if (chr_len[c] < 3) psx_throw(PSX_EINVAL, "chromosome %u is shorter than 3 bases", c);
        g->woff[c] = w;
        g->boff[c + 1] = g->boff[c] + chr_len[c];
        w += ((uint64_t)chr_len[c] + 15u) / 16u;
    }
    g->words = w;
}

extern "C" uint64_t psx_genome_words(uint32_t n_chr, const uint32_t* chr_len) {
    uint64_t w = 0;
    for (uint32_t c = 0; c < n_chr; ++c) w += ((uint64_t)chr_len[c] + 15u) / 16u;
    return w;
}

extern "C" psx_status psx_genome_synth(psx_ctx* ctx, uint32_t n_chr, const uint32_t* chr_len,
                                       uint64_t ref_seed, psx_genome** out) {
    if (!ctx || !chr_len || !out) return PSX_EINVAL;
    psx_genome* g = new psx_genome();
    g->ctx = ctx;
    psx_status st = psx_guard(ctx, [&] {
        genome_layout(g, n_chr, chr_len);
        PSX_HIP(hipSetDevice(ctx->device));
        g->packed.alloc(g->words * 4u);
        genome_synth_args a{};
        a.packed = g->packed.as<uint32_t>();
        a.seed = ref_seed;
        a.n_chr = n_chr;
        a.qoff[0] = 0;
        for (uint32_t c = 0; c < n_chr; ++c) {
            a.len[c] = chr_len[c];
            a.woff[c] = g->woff[c];
            a.qoff[c + 1] = a.qoff[c] + (((uint64_t)chr_len[c] + 15u) / 16u + 3u) / 4u;
        }
        const uint64_t blocks = (a.qoff[n_chr] + 255u) / 256u;
        hipLaunchKernelGGL(k_genome_synth, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream, a);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (st != PSX_OK) { delete g; return st; }
    *out = g;
    return PSX_OK;
}

extern "C" psx_status psx_genome_upload(psx_ctx* ctx, uint32_t n_chr, const uint32_t* chr_len,
                                        const uint32_t* packed, psx_genome** out) {
    if (!ctx || !chr_len || !packed || !out) return PSX_EINVAL;
    psx_genome* g = new psx_genome();
    g->ctx = ctx;
    psx_status st = psx_guard(ctx, [&] {
        genome_layout(g, n_chr, chr_len);
        PSX_HIP(hipSetDevice(ctx->device));
        g->packed.alloc(g->words * 4u);
        PSX_HIP(hipMemcpyAsync(g->packed.p, packed, g->words * 4u, hipMemcpyHostToDevice, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
    if (st != PSX_OK) { delete g; return st; }
    *out = g;
    return PSX_OK;
}

extern "C" psx_status psx_genome_download(psx_ctx* ctx, const psx_genome* g, uint32_t* packed,
                                          uint64_t n_words) {
    if (!ctx || !g || !packed) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        if (n_words < g->words) psx_throw(PSX_EINVAL, "buffer holds %llu words, need %llu",
                                          (unsigned long long)n_words, (unsigned long long)g->words);
        PSX_HIP(hipSetDevice(ctx->device));
        PSX_HIP(hipMemcpyAsync(packed, g->packed.p, g->words * 4u, hipMemcpyDeviceToHost, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" void psx_genome_free(psx_genome* g) {
    if (!g) return;
    if (g->ctx) (void)hipSetDevice(g->ctx->device);
    delete g;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t psx_mutations_bytes(uint64_t n) {
    return align256(n * sizeof(psx_mutation)) + 4 * align256(n * 4u) + align256(8 * (PSX_MAX_CHROMOSOMES + 1));
}

void psx_mutations_carve(psx_mutations* m, const std::shared_ptr<dbuf>& block, size_t off) {
    uint8_t* b = (uint8_t*)block->p + off;
    const uint64_t n = m->n;
    m->store = block;
    m->recs = (psx_mutation*)b;
    b += align256(n * sizeof(psx_mutation));
    m->pos = (uint32_t*)b;
    b += align256(n * 4u);
    m->lo = (uint32_t*)b;
    b += align256(n * 4u);
    m->hi = (uint32_t*)b;
    b += align256(n * 4u);
    m->info = (uint32_t*)b;
    b += align256(n * 4u);
    m->d_chr_begin = (uint64_t*)b;
}

// m->recs must hold m->n records in SID order with sid_head set; fills the
// SoA arrays, the chromosome ranges and the per-stratum index.
void psx_mutations_finish(psx_ctx* ctx, psx_mutations* m, const std::vector<uint64_t>& chr_begin) {
    m->chr_begin = chr_begin;
    m->chr_begin.resize(PSX_MAX_CHROMOSOMES + 1, m->n);
    PSX_HIP(hipMemcpyAsync(m->d_chr_begin, m->chr_begin.data(), 8 * (PSX_MAX_CHROMOSOMES + 1),
                           hipMemcpyHostToDevice, ctx->stream));
    psx_mutations_finish_device(ctx, m, true);
}

void psx_mutations_finish_device(psx_ctx* ctx, psx_mutations* m, bool sync) {
    const uint64_t n = m->n;
    if (n) {
        hipLaunchKernelGGL(k_mut_soa, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, ctx->stream, m->recs, n,
                           m->pos, m->lo, m->hi, m->info);
        PSX_HIP(hipGetLastError());
    }
    psx_mutations_index(ctx, m, m->chr_len, sync);
}

const psx_mut_index* psx_mutations_find_index(psx_mutations* m, const std::vector<uint32_t>& chr_len) {
    std::lock_guard<std::mutex> g(m->idx_mu);
    for (const auto& ix : m->index)
        if (ix->chr_len == chr_len) return ix.get();
    return nullptr;
}

std::vector<uint64_t> psx_index_layout(const std::vector<uint32_t>& chr_len) {
    std::vector<uint64_t> off(chr_len.size() + 1, 0);
    for (size_t c = 0; c < chr_len.size(); ++c) off[c + 1] = off[c] + (chr_len[c] + PSX_STRATUM - 1) / PSX_STRATUM + 1;
    return off;
}

const psx_mut_index& psx_mutations_index(psx_ctx* ctx, psx_mutations* m, const std::vector<uint32_t>& chr_len,
                                         bool sync) {
    const uint32_t n_chr = (uint32_t)chr_len.size();
    if (n_chr == 0 || n_chr > PSX_MAX_CHROMOSOMES) psx_throw(PSX_EINVAL, "invalid number of chromosomes");
    std::lock_guard<std::mutex> g(m->idx_mu);
    for (const auto& ix : m->index)
        if (ix->chr_len == chr_len) return *ix;
    auto ix = std::make_unique<psx_mut_index>();
    ix->chr_len = chr_len;
    bucket_args a{};
    a.pos = m->pos;
    a.n_chr = n_chr;
    ix->bucket_off = psx_index_layout(chr_len);
    a.chr_begin = m->d_chr_begin;
    for (uint32_t c = 0; c <= n_chr; ++c) a.boff[c] = ix->bucket_off[c];
    ix->store = std::make_shared<dbuf>();
    ix->store->alloc(ix->bucket_off[n_chr] * 4u);
    ix->bucket = ix->store->as<uint32_t>();
    a.bucket = ix->bucket;
    hipLaunchKernelGGL(k_mut_bucket, dim3((uint32_t)((ix->bucket_off[n_chr] + 255) / 256)), dim3(256), 0,
                       ctx->stream, a);
    PSX_HIP(hipGetLastError());
    // callers on other streams (serving lanes) read the index without waiting on ctx's stream; a caller
    // passing sync = false synchronises ctx's stream itself before the table is used
    if (sync) PSX_HIP(hipStreamSynchronize(ctx->stream));
    m->index.push_back(std::move(ix));
    return *m->index.back();
}

// Host path (replay uploads): SID sort with ties in input order.
void psx_mutations_build(psx_ctx* ctx, psx_mutations* m, std::vector<psx_mutation>& recs,
                         const std::vector<uint32_t>& chr_len) {
    const uint64_t n = recs.size();
    std::vector<uint64_t> key(n), idx(n);
    for (uint64_t i = 0; i < n; ++i) key[i] = psx_sid_key(recs[i]);
    std::iota(idx.begin(), idx.end(), 0ull);
    std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return key[a] < key[b]; });
    std::vector<psx_mutation> sorted(n);
    for (uint64_t i = 0; i < n; ++i) {
        sorted[i] = recs[idx[i]];
        sorted[i].sid_head = (i == 0 || key[idx[i]] != key[idx[i - 1]]) ? 1 : 0;
    }
    recs.swap(sorted);
    std::vector<uint64_t> chr_begin(PSX_MAX_CHROMOSOMES + 1, n);
    {
        uint64_t i = 0;
        for (uint32_t c = 0; c <= PSX_MAX_CHROMOSOMES; ++c) {
            while (i < n && recs[i].chr < c) ++i;
            chr_begin[c] = i;
        }
    }
    m->n = n;
    m->chr_len = chr_len;
    auto block = std::make_shared<dbuf>();
    block->alloc(psx_mutations_bytes(n));
    psx_mutations_carve(m, block, 0);
    if (n)
        PSX_HIP(hipMemcpyAsync(m->recs, recs.data(), n * sizeof(psx_mutation), hipMemcpyHostToDevice,
                               ctx->stream));
    psx_mutations_finish(ctx, m, chr_begin);
    PSX_HIP(hipStreamSynchronize(ctx->stream));
}

extern "C" psx_status psx_mutations_upload(psx_ctx* ctx, const psx_forest* f, const psx_mutation* recs,
                                           uint64_t n, uint32_t ploidy, psx_mutations** out) {
    if (!ctx || !f || (!recs && n) || !out) return PSX_EINVAL;
    psx_mutations* m = new psx_mutations();
    m->ctx = ctx;
    psx_status st = psx_guard(ctx, [&] {
        if (ploidy < 1 || ploidy > 255) psx_throw(PSX_EINVAL, "invalid ploidy %u", ploidy);
        std::vector<psx_mutation> v(recs, recs + n);
        uint32_t max_chr = 0;
        for (const auto& r : v) {
            if (r.chr >= PSX_MAX_CHROMOSOMES) psx_throw(PSX_EINVAL, "chromosome index %u out of range", r.chr);
            if (r.allele >= ploidy) psx_throw(PSX_EINVAL, "allele %u out of range", r.allele);
            if (r.kind > PSX_INS || r.nature > PSX_GERMINAL || r.ref > 3 || r.alt > 3 || r.len > 15)
                psx_throw(PSX_EINVAL, "malformed mutation record");
            if (r.leaf_lo > r.leaf_hi || r.leaf_hi > f->n_leaves)
                psx_throw(PSX_EINVAL, "leaf range [%u, %u) outside the forest", r.leaf_lo, r.leaf_hi);
            max_chr = std::max(max_chr, r.chr + 1);
        }
        m->ploidy = ploidy;
        m->n_leaves = f->n_leaves;
        // chromosome lengths are only known from the genome at sequencing time;
        // index every chromosome up to the largest position seen
        std::vector<uint32_t> chr_len(std::max(max_chr, 1u), 0);
        for (const auto& r : v) chr_len[r.chr] = std::max(chr_len[r.chr], r.pos + 1);
        PSX_HIP(hipSetDevice(ctx->device));
        psx_mutations_build(ctx, m, v, chr_len);
    });
    if (st != PSX_OK) { delete m; return st; }
    *out = m;
    return PSX_OK;
}

extern "C" psx_status psx_mutations_count(const psx_mutations* m, uint64_t* n) {
    if (!m || !n) return PSX_EINVAL;
    *n = m->n;
    return PSX_OK;
}

void psx_cna_check(const std::vector<psx_cna_event>& ev, uint32_t n_leaves, uint32_t ploidy,
                   const std::vector<uint32_t>& chr_len) {
    if (!ev.empty() && ploidy > PSX_MAX_CNA_PLOIDY)
        psx_throw(PSX_EINVAL, "copy-number events need a ploidy of at most %u", PSX_MAX_CNA_PLOIDY);
    for (size_t i = 0; i < ev.size(); ++i) {
        const psx_cna_event& e = ev[i];
        const bool wgd = e.kind == PSX_CNA_WGD;
        if (e.kind > PSX_CNA_WGD || e.leaf_lo > e.leaf_hi || e.leaf_hi > n_leaves ||
            (!wgd && (e.chr >= chr_len.size() || e.start >= e.end || e.end > chr_len[e.chr] || e.allele >= ploidy ||
                      (e.kind == PSX_CNA_AMP && e.copies == 0))) ||
            (wgd && e.chr != PSX_ALL_CHROMOSOMES) || (i && ev[i - 1].node > e.node))
            psx_throw(PSX_EINVAL, "invalid copy-number event %zu", i);
    }
}

extern "C" psx_status psx_mutations_cna(const psx_mutations* m, psx_cna_event* ev, uint32_t cap, uint32_t* n) {
    if (!m || !n) return PSX_EINVAL;
    *n = (uint32_t)m->cna.size();
    if (ev)
        for (uint32_t i = 0; i < cap && i < *n; ++i) ev[i] = m->cna[i];
    return PSX_OK;
}

extern "C" psx_status psx_mutations_set_cna(psx_ctx* ctx, psx_mutations* m, const psx_cna_event* ev, uint32_t n) {
    if (!ctx || !m || (!ev && n)) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        std::vector<psx_cna_event> v(ev, ev + n);
        psx_cna_check(v, m->n_leaves, m->ploidy, m->chr_len);
        m->cna = std::move(v);
    });
}

extern "C" psx_status psx_mutations_download(psx_ctx* ctx, const psx_mutations* m, psx_mutation* recs,
                                             uint64_t n) {
    if (!ctx || !m || (!recs && m->n)) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        if (n < m->n) psx_throw(PSX_EINVAL, "buffer holds %llu records, need %llu", (unsigned long long)n,
                                (unsigned long long)m->n);
        PSX_HIP(hipSetDevice(ctx->device));
        if (m->n)
            PSX_HIP(hipMemcpyAsync(recs, m->recs, m->n * sizeof(psx_mutation), hipMemcpyDeviceToHost,
                                   ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" void psx_mutations_free(psx_mutations* m) {
    if (!m) return;
    if (m->ctx) (void)hipSetDevice(m->ctx->device);
    delete m;
}
