// SBS context index of a genome's reference on gfx950: the reference engine's
// context index (build_contex_index, src/mutation_engine.cpp:213-261, a RACES
// ContextIndex with context_sampling), restated. A site is a position q with
// 1 <= q < len - 1; its context is the pyrimidine-centred trinucleotide read
// on the strand where b[q] is C or T (32 contexts). Per context, every s-th
// site in genome order from the first is kept as a 32-bit global position.
//
//   k_ctx_scan<false>  one wave per tile of CTX_TILE positions: the tile's
//                      sites per context -> cnt[ctx][tile]
//   (scan)             exclusive sum of cnt in [ctx][tile] order
//   k_ctx_bounds       the 33 context boundaries of that sum
//   k_ctx_scan<true>   the same waves: each site's rank within its context
//                      (one ballot per distinct context among the 64 lanes)
//                      and every s-th site written to its context's list
//
// The lists come out in genome order with no sort, so the build is
// deterministic. The work is integer and bounded by VALU issue (about 32
// ballots per 64 positions per pass); it runs once per genome and sampling.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "psx_internal.h"

namespace {

constexpr uint32_t CTX_TILE = 4096;  // positions per wave

struct ctx_args {
    const uint32_t* packed;
    const uint64_t* woff;
    const uint32_t* len;
    const uint64_t* boff;
    const uint32_t* tile_chr;
    const uint32_t* tile_first;
    uint32_t n_tiles;
    uint32_t sampling;
    uint32_t* cnt;          // [32][n_tiles] (count pass)
    const uint32_t* scan;   // exclusive sum of cnt (list pass)
    const uint64_t* off;    // [33] list ranges (list pass)
    uint32_t* list;
    uint32_t strand_bit;    // list entries carry the site's strand in bit 31
};

__device__ __forceinline__ uint32_t wbase(const uint32_t* words, uint32_t i) {
    return (words[i >> 4] >> (2u * (i & 15u))) & 3u;
}

// A = 0, C = 1, G = 2, T = 3: C and T are the pyrimidines
__device__ __forceinline__ uint32_t ctx_of(uint32_t b5, uint32_t m, uint32_t b3) {
    const bool pyr = (m & 1u) != 0u;
    const uint32_t mid = pyr ? m : 3u - m, five = pyr ? b5 : 3u - b3, three = pyr ? b3 : 3u - b5;
    return ((mid == 3u ? 1u : 0u) << 4) | (five << 2) | three;
}

template <bool LIST>
__global__ void __launch_bounds__(256) k_ctx_scan(const ctx_args a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (tile >= a.n_tiles) return;  // the whole wave
    const uint32_t c = a.tile_chr[tile], first = a.tile_first[tile];
    const uint32_t len = a.len[c];
    const uint32_t end = min(first + CTX_TILE, len);
    const uint32_t* words = a.packed + a.woff[c];
    const uint32_t g0 = (uint32_t)a.boff[c];
    const uint64_t lt = (1ull << lane) - 1ull;
    // lane k < 32: sites of context k so far in the tile; in the list pass
    // plus the context's sites in the tiles before this one
    uint32_t run = 0;
    if (LIST && lane < 32u) run = a.scan[lane * a.n_tiles + tile] - a.scan[lane * a.n_tiles];
    for (uint32_t q0 = first; q0 < end; q0 += 64u) {
        const uint32_t q = q0 + lane;
        uint32_t k = 32u, purine = 0u;
        if (q >= 1u && q + 1u < len && q < end) {
            const uint32_t m = wbase(words, q);
            k = ctx_of(wbase(words, q - 1u), m, wbase(words, q + 1u));
            purine = (m & 1u) ? 0u : 1u;
        }
        const uint32_t entry = (g0 + q) | ((purine & a.strand_bit) << 31);
        uint64_t pending = __ballot(k < 32u);
        while (pending) {
            const uint32_t kk = (uint32_t)__builtin_amdgcn_readlane((int)k, (int)__builtin_ctzll(pending));
            const uint64_t same = __ballot(k == kk);
            if (LIST) {
                const uint32_t r0 = (uint32_t)__builtin_amdgcn_readlane((int)run, (int)kk);
                if (k == kk) {
                    const uint32_t r = r0 + (uint32_t)__popcll(same & lt);
                    if (a.sampling == 1u) a.list[a.off[kk] + r] = entry;
                    else if (r % a.sampling == 0u) a.list[a.off[kk] + r / a.sampling] = entry;
                }
            }
            if (lane == kk) run += (uint32_t)__popcll(same);
            pending &= ~same;
        }
    }
    if (!LIST && lane < 32u) a.cnt[lane * a.n_tiles + tile] = run;
}

// out[k] = first flattened rank of context k (k < 32), out[32] = all sites
__global__ void k_ctx_bounds(const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ scan, uint32_t n_tiles,
                             uint32_t* __restrict__ out) {
    const uint32_t k = threadIdx.x;
    if (k < 32u) out[k] = scan[k * n_tiles];
    else if (k == 32u) {
        const size_t last = (size_t)32 * n_tiles - 1;
        out[32] = scan[last] + cnt[last];
    }
}

size_t align256(size_t n) { return (std::max<size_t>(n, 1) + 255) & ~size_t(255); }

}  // namespace

const psx_sbs_index& psx_genome_sbs_index(psx_ctx* ctx, psx_genome* g, uint32_t sampling) {
    if (sampling == 0) sampling = 1;
    std::lock_guard<std::mutex> lk(g->sbs_mu);
    auto it = g->sbs_index.find(sampling);
    if (it != g->sbs_index.end()) return *it->second;
    const uint64_t G = g->boff[g->n_chr];
    if (G > 0xFFFFFFFFull) psx_throw(PSX_EINVAL, "the reference is too long for the context index (at most 2^32 - 1 bases)");
    auto ix = std::make_unique<psx_sbs_index>();
    ix->sampling = sampling;
    hipStream_t s = ctx->stream;
    std::vector<uint32_t> tile_chr, tile_first;
    for (uint32_t c = 0; c < g->n_chr; ++c)
        for (uint64_t f = 0; f < g->len[c]; f += CTX_TILE) {
            tile_chr.push_back(c);
            tile_first.push_back((uint32_t)f);
        }
    const uint32_t n_tiles = (uint32_t)tile_chr.size();
    scratch_builder sb;
    const size_t o_len = sb.put_vec(g->len), o_woff = sb.put_vec(g->woff), o_boff = sb.put_vec(g->boff);
    const size_t o_tc = sb.put_vec(tile_chr), o_tf = sb.put_vec(tile_first);
    uint8_t* base = (uint8_t*)psx_scratch_upload(ctx, sb);
    const size_t n_cells = (size_t)32 * n_tiles;
    size_t scan_bytes = 0;
    PSX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (int)n_cells, s));
    dbuf tmp;
    const size_t t_cnt = 0, t_scan = align256(n_cells * 4), t_bounds = t_scan + align256(n_cells * 4);
    const size_t t_cub = t_bounds + align256(33 * 4);
    tmp.alloc(t_cub + align256(scan_bytes));
    uint8_t* tb = (uint8_t*)tmp.p;
    ctx_args a{};
    a.packed = g->packed.as<uint32_t>();
    a.len = (const uint32_t*)(base + o_len);
    a.woff = (const uint64_t*)(base + o_woff);
    a.boff = (const uint64_t*)(base + o_boff);
    a.tile_chr = (const uint32_t*)(base + o_tc);
    a.tile_first = (const uint32_t*)(base + o_tf);
    a.n_tiles = n_tiles;
    a.sampling = sampling;
    a.cnt = (uint32_t*)(tb + t_cnt);
    a.scan = (const uint32_t*)(tb + t_scan);
    const dim3 grid((n_tiles + 3) / 4);
    PSX_HIP(hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(k_ctx_scan<false>, grid, dim3(256), 0, s, a);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipcub::DeviceScan::ExclusiveSum(tb + t_cub, scan_bytes, a.cnt, (uint32_t*)(tb + t_scan), (int)n_cells, s));
    hipLaunchKernelGGL(k_ctx_bounds, dim3(1), dim3(64), 0, s, a.cnt, a.scan, n_tiles, (uint32_t*)(tb + t_bounds));
    PSX_HIP(hipGetLastError());
    uint32_t bounds[33];
    PSX_HIP(hipMemcpyAsync(bounds, tb + t_bounds, sizeof bounds, hipMemcpyDeviceToHost, s));
    PSX_HIP(hipStreamSynchronize(s));
    uint64_t total = 0;
    for (uint32_t k = 0; k < 32; ++k) {
        ix->count[k] = bounds[k + 1] - bounds[k];
        ix->kept[k] = (ix->count[k] + sampling - 1) / sampling;
        ix->off[k] = total;
        total += ix->kept[k];
    }
    ix->off[32] = total;
    ix->d_list.alloc(std::max<uint64_t>(total, 1) * 4);
    ix->d_meta.alloc(65 * 8);
    uint64_t meta[65];
    std::copy(ix->kept, ix->kept + 32, meta);
    std::copy(ix->off, ix->off + 33, meta + 32);
    PSX_HIP(hipMemcpyAsync(ix->d_meta.p, meta, sizeof meta, hipMemcpyHostToDevice, s));
    a.off = ix->d_meta.as<uint64_t>() + 32;
    a.list = ix->d_list.as<uint32_t>();
    ix->strand_bit = g->boff[g->n_chr] <= (uint64_t(1) << 31);
    a.strand_bit = ix->strand_bit ? 1u : 0u;
    if (total) {
        hipLaunchKernelGGL(k_ctx_scan<true>, grid, dim3(256), 0, s, a);
        PSX_HIP(hipGetLastError());
    }
    PSX_HIP(hipEventRecord(ctx->ev1, s));
    PSX_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ix->build_ms = ms;
    auto& slot = g->sbs_index[sampling];
    slot = std::move(ix);
    return *slot;
}

extern "C" psx_status psx_genome_context_index(psx_ctx* ctx, psx_genome* g, uint32_t sampling, uint64_t* counts,
                                               uint64_t* kept, uint32_t* list, uint64_t list_cap, double* build_ms) {
    if (!ctx || !g) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        const psx_sbs_index& ix = psx_genome_sbs_index(ctx, g, sampling);
        if (counts) std::copy(ix.count, ix.count + 32, counts);
        if (kept) std::copy(ix.kept, ix.kept + 32, kept);
        if (build_ms) *build_ms = ix.build_ms;
        if (list) {
            if (list_cap < ix.off[32]) psx_throw(PSX_EINVAL, "the list needs %llu entries, the buffer holds %llu",
                                                 (unsigned long long)ix.off[32], (unsigned long long)list_cap);
            if (ix.off[32])
                PSX_HIP(hipMemcpy(list, ix.d_list.p, ix.off[32] * 4, hipMemcpyDeviceToHost));
            if (ix.strand_bit)  // positions only
                for (uint64_t i = 0; i < ix.off[32]; ++i) list[i] &= 0x7FFFFFFFu;
        }
    });
}
