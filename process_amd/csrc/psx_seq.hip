// simulate_seq on gfx950: the ReadSimulator call of the reference
// (src/seq_simulation.cpp:354-429, RACES call at :371-372) in count-only
// mode, its read-materialising analogue (write_SAM = TRUE, :564) and the
// batched form.
//
// Work decomposition. A segment is one (job, output sample, chromosome); its
// W = len - L + 1 template starts are cut into strata of PSX_STRATUM starts,
// and stratum j owns a fixed number of templates (part of the sampling model,
// so results do not depend on the launch). A work item is a run of strata of
// one segment; persistent blocks pull items from a device counter.
//
// Per stratum a block stages the mutation records whose positions the reads
// can reach (pos, leaf range, packed info) in LDS, with a 64-base bucket index,
// and keeps LDS alt/depth counters; reads are generated from counters, looked
// up in LDS, and the block flushes its counters with one global atomic per
// touched record. All counting is integer.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "psx_internal.h"

namespace {

constexpr uint32_t NB_SHIFT = 6;      // 64-base LDS buckets
constexpr uint32_t MAX_READ = 4096;   // read size and insert reach limit
constexpr uint32_t NB_MAX = (PSX_STRATUM + MAX_READ) / 64 + 2;

struct job_dev {
    uint64_t seed;
    const uint32_t* pos;
    const uint32_t* lo;
    const uint32_t* hi;
    const uint32_t* info;
    const uint32_t* bucket;
    const uint32_t* sample_leaves;
    int32_t* alt;
    int32_t* depth;
    uint8_t* present;
    const uint64_t* ins_T;
    const uint64_t* err_T;
    const uint32_t* packed;      // reference (read materialisation)
    uint64_t n_mut;
    uint64_t thr_normal;
    uint32_t ins_k0, ins_n, err_n;
    uint32_t L, pe, ploidy, n_leaves, preneo, reach;
    uint32_t n_out;
    uint32_t purity_lt1;
};

struct seg_dev {
    uint64_t T;             // templates of the segment
    uint64_t out_off;       // so * n_mut
    uint64_t read_base;     // first read index of the segment (materialisation)
    uint64_t chr_woff;      // first packed word of the chromosome
    uint32_t job, so, chr, len;
    uint32_t W, J, bucket_off, n_bucket;
    uint32_t leaf_off, n_cells;  // n_cells == 0: normal sample
};

struct item_dev {
    uint32_t seg, j0, j1, pad;
};

struct pres_dev {  // one per (job, output sample) for the init kernel
    uint64_t begin;        // global index of the first (so, m) entry
    uint32_t job, so, leaf_off, n_cells;
};

struct seq_launch {
    const job_dev* jobs;
    const seg_dev* segs;
    const item_dev* items;
    uint64_t n_items;
    unsigned long long* queue;
    uint32_t* reads_out;     // materialisation only
    uint32_t read_words;
};

// ------------------------------------------------------------- record views

struct lds_view {
    const uint32_t* pos;
    const uint32_t* lo;
    const uint32_t* hi;
    const uint32_t* info;
    uint32_t* alt;
    uint32_t* dep;
    const uint16_t* bucket;
    uint32_t n, b;
    __device__ __forceinline__ uint32_t first(uint32_t x) const {
        uint32_t i = bucket[(x - b) >> NB_SHIFT];
        while (i < n && pos[i] < x) ++i;
        return i;
    }
    __device__ __forceinline__ void add_dep(uint32_t q) const { atomicAdd(&dep[q], 1u); }
    __device__ __forceinline__ void add_alt(uint32_t q) const { atomicAdd(&alt[q], 1u); }
};

struct glb_view {  // window larger than the LDS capacity
    const uint32_t* pos;
    const uint32_t* lo;
    const uint32_t* hi;
    const uint32_t* info;
    int32_t* alt;
    int32_t* dep;
    uint32_t n;
    __device__ __forceinline__ uint32_t first(uint32_t x) const {
        uint32_t a = 0, z = n;
        while (a < z) {
            const uint32_t mid = (a + z) >> 1;
            if (pos[mid] < x) a = mid + 1; else z = mid;
        }
        return a;
    }
    __device__ __forceinline__ void add_dep(uint32_t q) const { atomicAdd(&dep[q], 1); }
    __device__ __forceinline__ void add_alt(uint32_t q) const { atomicAdd(&alt[q], 1); }
};

struct read_src {
    uint32_t leaf, allele;
};

__device__ __forceinline__ bool carried(const job_dev& J, uint32_t info, uint32_t lo, uint32_t hi, read_src s) {
    if (PSX_INFO_ALLELE(info) != s.allele) return false;
    if (s.leaf == J.n_leaves) {
        const uint32_t nat = PSX_INFO_NATURE(info);
        return nat == PSX_GERMINAL || (nat == PSX_PRENEOPLASTIC && J.preneo);
    }
    return lo <= s.leaf && s.leaf < hi;
}

// Error set of one read (oracle: error_set): K ~ Binomial(L, e) from word 0,
// then K distinct offsets, one word each; a repeated offset takes the next
// word. Slots hold offset | substitution << 12.
__device__ uint32_t error_set(const job_dev& J, uint32_t r, uint32_t j, uint32_t c2, uint32_t mate,
                              uint16_t* slots, uint32_t stride) {
    const uint32_t base = (PSX_P_ERROR << 24) | (mate << 16);
    psx_u4 w = psx_draw(J.seed, r, j, c2, base);
    const uint32_t K = psx_cdf_pick(J.err_T, J.err_n, w.x);
    uint32_t cnt = 0;
    for (uint32_t t = 1; cnt < K && t < 1 + 4 * PSX_MAX_READ_ERRORS; ++t) {
        if ((t & 3u) == 0) w = psx_draw(J.seed, r, j, c2, base | (t >> 2));
        const uint32_t word = psx_word(w, (int)(t & 3u));
        const uint32_t o = psx_mulhi32(word, J.L);
        bool dup = false;
        for (uint32_t q = 0; q < cnt; ++q) dup |= (slots[q * stride] & 0xFFFu) == o;
        if (dup) continue;
        slots[cnt * stride] = (uint16_t)(o | ((((word & 0xFFFFu) * 3u) >> 16) << 12));
        ++cnt;
    }
    return cnt;
}

// One read [x, x + L): depth for every record at a covered position, an
// alternative read for each SID head whose alternative is observed (SNV) or
// carried (indel). Matches oracle seq_read.
template <class V>
__device__ void count_read(const job_dev& J, const V& v, uint32_t x, read_src src, uint32_t r, uint32_t j,
                           uint32_t c2, uint32_t mate, uint16_t* slots, uint32_t stride) {
    uint32_t i = v.first(x);
    const uint32_t end = x + J.L;
    if (i >= v.n || v.pos[i] >= end) return;
    int ne = -1;
    while (i < v.n && v.pos[i] < end) {
        const uint32_t p = v.pos[i];
        const uint32_t g0 = i;
        uint32_t g1 = i + 1;
        while (g1 < v.n && v.pos[g1] == p) ++g1;
        uint32_t tb = PSX_INFO_REF(v.info[g0]);
        for (uint32_t q = g0; q < g1; ++q) {
            const uint32_t in = v.info[q];
            if (PSX_INFO_KIND(in) == PSX_SNV && carried(J, in, v.lo[q], v.hi[q], src)) {
                tb = PSX_INFO_ALT(in);
                break;
            }
        }
        uint32_t ob = tb;
        if (J.err_n) {
            if (ne < 0) ne = (int)error_set(J, r, j, c2, mate, slots, stride);
            for (int q = 0; q < ne; ++q) {
                const uint32_t s = slots[q * stride];
                if ((s & 0xFFFu) == p - x) {
                    const uint32_t sub = s >> 12;
                    ob = sub < tb ? sub : sub + 1;
                }
            }
        }
        for (uint32_t q = g0; q < g1; ++q) {
            v.add_dep(q);
            const uint32_t in = v.info[q];
            if (!PSX_INFO_HEAD(in)) continue;
            if (PSX_INFO_KIND(in) == PSX_SNV) {
                if (ob == PSX_INFO_ALT(in)) v.add_alt(q);
            } else {
                bool car = false;
                for (uint32_t q2 = q; q2 < g1 && (q2 == q || !PSX_INFO_HEAD(v.info[q2])); ++q2)
                    car |= carried(J, v.info[q2], v.lo[q2], v.hi[q2], src);
                if (car) v.add_alt(q);
            }
        }
        i = g1;
    }
}

struct tmpl {
    uint32_t x, x2;
    read_src src;
};

// Template r of stratum j: start, source (normal with probability 1 - purity;
// else a uniform (cell, copy) of the sample), insert size for paired ends.
__device__ __forceinline__ tmpl make_template(const job_dev& J, const seg_dev& S, uint32_t r, uint32_t j,
                                              uint32_t b, uint32_t e) {
    const uint32_t c2 = (S.chr << 16) | S.so;
    const psx_u4 w = psx_draw(J.seed, r, j, c2, PSX_P_TEMPLATE << 24);
    tmpl t;
    t.x = b + psx_uniform32(w.x, e - b);
    if (S.n_cells == 0 || (uint64_t)w.y < J.thr_normal) {
        t.src.leaf = J.n_leaves;
        t.src.allele = psx_uniform32(w.z, J.ploidy);
    } else {
        const uint32_t X = psx_uniform32(w.z, S.n_cells * J.ploidy);
        uint32_t cell, allele;
        if (J.ploidy == 2) { cell = X >> 1; allele = X & 1u; }
        else { cell = X / J.ploidy; allele = X - cell * J.ploidy; }
        t.src.leaf = J.sample_leaves[S.leaf_off + cell];
        t.src.allele = allele;
    }
    t.x2 = t.x;
    if (J.pe) {
        uint32_t I = J.ins_k0 + psx_cdf_pick(J.ins_T, J.ins_n, w.w);
        if (I < J.L) I = J.L;
        uint64_t x2 = (uint64_t)t.x + I - J.L;
        if (x2 > S.len - J.L) x2 = S.len - J.L;
        t.x2 = (uint32_t)x2;
    }
    return t;
}

// ------------------------------------------------------------ count kernel

template <int BLOCK, int CAP>
__global__ void __launch_bounds__(BLOCK) k_seq_count(const seq_launch L) {
    __shared__ uint32_t s_pos[CAP], s_lo[CAP], s_hi[CAP], s_info[CAP], s_alt[CAP], s_dep[CAP];
    __shared__ uint16_t s_bucket[NB_MAX];
    __shared__ uint16_t s_err[PSX_MAX_READ_ERRORS * BLOCK];
    __shared__ unsigned long long s_item;
    const uint32_t tid = threadIdx.x;
    uint16_t* slots = s_err + tid;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(L.queue, 1ull);
        __syncthreads();
        const unsigned long long item = s_item;
        __syncthreads();
        if (item >= L.n_items) break;
        const item_dev it = L.items[item];
        const seg_dev S = L.segs[it.seg];
        const job_dev J = L.jobs[S.job];
        const uint32_t c2 = (S.chr << 16) | S.so;
        for (uint32_t j = it.j0; j < it.j1; ++j) {
            const uint32_t b = j * PSX_STRATUM;
            const uint32_t e = min(b + PSX_STRATUM, S.W);
            const uint64_t nj = S.T * e / S.W - S.T * b / S.W;
            const uint32_t wend = min(S.len, e - 1 + J.reach);
            const uint32_t k0 = J.bucket[S.bucket_off + j];
            const uint32_t kb = min((wend + PSX_STRATUM - 1) / PSX_STRATUM, S.n_bucket - 1);
            const uint32_t k1 = J.bucket[S.bucket_off + kb];
            const uint32_t n_win = k1 - k0;
            if (n_win <= (uint32_t)CAP) {
                for (uint32_t i = tid; i < n_win; i += BLOCK) {
                    s_pos[i] = J.pos[k0 + i];
                    s_lo[i] = J.lo[k0 + i];
                    s_hi[i] = J.hi[k0 + i];
                    s_info[i] = J.info[k0 + i];
                    s_alt[i] = 0;
                    s_dep[i] = 0;
                }
                __syncthreads();
                const uint32_t nb = ((wend - b) >> NB_SHIFT) + 1;
                for (uint32_t q = tid; q < nb; q += BLOCK) {
                    const uint32_t target = b + (q << NB_SHIFT);
                    uint32_t a = 0, z = n_win;
                    while (a < z) {
                        const uint32_t mid = (a + z) >> 1;
                        if (s_pos[mid] < target) a = mid + 1; else z = mid;
                    }
                    s_bucket[q] = (uint16_t)a;
                }
                __syncthreads();
                const lds_view v{s_pos, s_lo, s_hi, s_info, s_alt, s_dep, s_bucket, n_win, b};
                for (uint32_t r = tid; r < (uint32_t)nj; r += BLOCK) {
                    const tmpl t = make_template(J, S, r, j, b, e);
                    count_read(J, v, t.x, t.src, r, j, c2, 0u, slots, BLOCK);
                    if (J.pe) count_read(J, v, t.x2, t.src, r, j, c2, 1u, slots, BLOCK);
                }
                __syncthreads();
                int32_t* ga = J.alt + S.out_off + k0;
                int32_t* gd = J.depth + S.out_off + k0;
                for (uint32_t i = tid; i < n_win; i += BLOCK) {
                    const uint32_t d = s_dep[i];
                    if (d) {
                        atomicAdd(&gd[i], (int32_t)d);
                        const uint32_t a = s_alt[i];
                        if (a) atomicAdd(&ga[i], (int32_t)a);
                    }
                }
                __syncthreads();
            } else {
                const glb_view v{J.pos + k0, J.lo + k0, J.hi + k0, J.info + k0, J.alt + S.out_off + k0,
                                 J.depth + S.out_off + k0, n_win};
                for (uint32_t r = tid; r < (uint32_t)nj; r += BLOCK) {
                    const tmpl t = make_template(J, S, r, j, b, e);
                    count_read(J, v, t.x, t.src, r, j, c2, 0u, slots, BLOCK);
                    if (J.pe) count_read(J, v, t.x2, t.src, r, j, c2, 1u, slots, BLOCK);
                }
            }
        }
    }
}

// Presence (the sample holds the mutation), and zeroed counters.
__global__ void __launch_bounds__(256) k_seq_init(const job_dev* __restrict__ jobs, const pres_dev* __restrict__ ps,
                                                  uint32_t n_ps, uint64_t total,
                                                  unsigned long long* __restrict__ queue) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) *queue = 0ull;
    if (q >= total) return;
    uint32_t a = 0, z = n_ps;
    while (a < z) {
        const uint32_t mid = (a + z) >> 1;
        if (ps[mid].begin <= q) a = mid + 1; else z = mid;
    }
    const pres_dev P = ps[a - 1];
    const job_dev& J = jobs[P.job];
    const uint64_t m = q - P.begin;
    const uint64_t o = (uint64_t)P.so * J.n_mut + m;
    J.alt[o] = 0;
    J.depth[o] = 0;
    const uint32_t info = J.info[m];
    const uint32_t nat = PSX_INFO_NATURE(info);
    const bool normal_has = nat == PSX_GERMINAL || (nat == PSX_PRENEOPLASTIC && J.preneo);
    bool pres;
    if (P.n_cells == 0) {
        pres = normal_has;
    } else {
        pres = J.purity_lt1 && normal_has;
        if (!pres) {
            const uint32_t lo = J.lo[m], hi = J.hi[m];
            const uint32_t* sl = J.sample_leaves + P.leaf_off;
            uint32_t x = 0, y = P.n_cells;
            while (x < y) {
                const uint32_t mid = (x + y) >> 1;
                if (sl[mid] < lo) x = mid + 1; else y = mid;
            }
            pres = x < P.n_cells && sl[x] < hi;
        }
    }
    J.present[o] = pres ? 1 : 0;
}

// --------------------------------------------------- read materialisation

__device__ __forceinline__ uint32_t ref_word(const uint32_t* chr_words, uint32_t n_words, uint32_t x) {
    // 16 bases starting at x (bases past the chromosome end read as A)
    const uint32_t w = x >> 4, sh = (x & 15u) * 2u;
    const uint32_t lo = chr_words[w];
    if (sh == 0) return lo;
    const uint32_t hi = w + 1 < n_words ? chr_words[w + 1] : 0u;
    return (lo >> sh) | (hi << (32u - sh));
}

// Writes one read: header word 0 = start, word 1 = leaf | allele << 26 |
// mate << 30 | (errors > 0) << 31, then ceil(L/16) words of 2-bit bases with
// carried SNVs and the read's errors applied.
template <class V>
__device__ void write_read(const job_dev& J, const seg_dev& S, const V& v, uint32_t x, read_src src,
                           uint32_t r, uint32_t j, uint32_t c2, uint32_t mate, uint16_t* slots, uint32_t stride,
                           uint32_t* out, uint32_t out_stride) {
    const uint32_t* cw = J.packed + S.chr_woff;
    const uint32_t chr_words = (S.len + 15u) >> 4;
    const uint32_t nw = (J.L + 15u) >> 4;
    const uint32_t last_bits = (J.L & 15u) ? (J.L & 15u) * 2u : 32u;
    uint32_t ne = 0;
    if (J.err_n) ne = error_set(J, r, j, c2, mate, slots, stride);
    for (uint32_t k = 0; k < nw; ++k) {
        const uint32_t x0 = x + (k << 4);
        uint32_t word = ref_word(cw, chr_words, x0);
        const uint32_t span = (k + 1 == nw) ? last_bits / 2u : 16u;
        if (span < 16u) word &= (1u << (2u * span)) - 1u;
        // carried SNVs in [x0, x0 + span): the first carried SNV at a position wins
        uint32_t i = v.first(x0);
        while (i < v.n && v.pos[i] < x0 + span) {
            const uint32_t p = v.pos[i];
            uint32_t g1 = i;
            bool done = false;
            while (g1 < v.n && v.pos[g1] == p) {
                const uint32_t in = v.info[g1];
                if (!done && PSX_INFO_KIND(in) == PSX_SNV && carried(J, in, v.lo[g1], v.hi[g1], src)) {
                    const uint32_t sh = (p - x0) * 2u;
                    word = (word & ~(3u << sh)) | (PSX_INFO_ALT(in) << sh);
                    done = true;
                }
                ++g1;
            }
            i = g1;
        }
        for (uint32_t q = 0; q < ne; ++q) {
            const uint32_t s = slots[q * stride];
            const uint32_t o = s & 0xFFFu;
            if (o >= (k << 4) && o < (k << 4) + span) {
                const uint32_t sh = (o - (k << 4)) * 2u;
                const uint32_t tb = (word >> sh) & 3u, sub = s >> 12;
                const uint32_t nb = sub < tb ? sub : sub + 1;
                word = (word & ~(3u << sh)) | (nb << sh);
            }
        }
        out[(2u + k) * out_stride] = word;
    }
    out[0] = x;
    out[out_stride] = src.leaf | (src.allele << 26) | (mate << 30) | ((ne > 0 ? 1u : 0u) << 31);
}

template <int BLOCK, int CAP>
__global__ void __launch_bounds__(BLOCK) k_seq_reads(const seq_launch L) {
    __shared__ uint32_t s_pos[CAP], s_lo[CAP], s_hi[CAP], s_info[CAP];
    __shared__ uint16_t s_bucket[NB_MAX];
    __shared__ uint16_t s_err[PSX_MAX_READ_ERRORS * BLOCK];
    __shared__ unsigned long long s_item;
    extern __shared__ uint32_t s_out[];  // [read word][BLOCK * reads per template]
    const uint32_t tid = threadIdx.x;
    uint16_t* slots = s_err + tid;
    const uint32_t RW = L.read_words;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(L.queue, 1ull);
        __syncthreads();
        const unsigned long long item = s_item;
        __syncthreads();
        if (item >= L.n_items) break;
        const item_dev it = L.items[item];
        const seg_dev S = L.segs[it.seg];
        const job_dev J = L.jobs[S.job];
        const uint32_t c2 = (S.chr << 16) | S.so;
        const uint32_t rpt = J.pe ? 2u : 1u;
        for (uint32_t j = it.j0; j < it.j1; ++j) {
            const uint32_t b = j * PSX_STRATUM;
            const uint32_t e = min(b + PSX_STRATUM, S.W);
            const uint64_t before = S.T * b / S.W;
            const uint64_t nj = S.T * e / S.W - before;
            const uint32_t wend = min(S.len, e - 1 + J.reach);
            const uint32_t k0 = J.bucket[S.bucket_off + j];
            const uint32_t kb = min((wend + PSX_STRATUM - 1) / PSX_STRATUM, S.n_bucket - 1);
            const uint32_t n_win = J.bucket[S.bucket_off + kb] - k0;
            const bool fits = n_win <= (uint32_t)CAP;
            if (fits) {
                for (uint32_t i = tid; i < n_win; i += BLOCK) {
                    s_pos[i] = J.pos[k0 + i];
                    s_lo[i] = J.lo[k0 + i];
                    s_hi[i] = J.hi[k0 + i];
                    s_info[i] = J.info[k0 + i];
                }
                __syncthreads();
                const uint32_t nb = ((wend - b) >> NB_SHIFT) + 1;
                for (uint32_t q = tid; q < nb; q += BLOCK) {
                    const uint32_t target = b + (q << NB_SHIFT);
                    uint32_t a = 0, z = n_win;
                    while (a < z) {
                        const uint32_t mid = (a + z) >> 1;
                        if (s_pos[mid] < target) a = mid + 1; else z = mid;
                    }
                    s_bucket[q] = (uint16_t)a;
                }
                __syncthreads();
            }
            const lds_view lv{s_pos, s_lo, s_hi, s_info, nullptr, nullptr, s_bucket, n_win, b};
            const glb_view gv{J.pos + k0, J.lo + k0, J.hi + k0, J.info + k0, nullptr, nullptr, n_win};
            const uint32_t stride = BLOCK * rpt;  // reads staged per batch
            for (uint32_t r0 = 0; r0 < (uint32_t)nj; r0 += BLOCK) {
                const uint32_t r = r0 + tid;
                if (r < (uint32_t)nj) {
                    const tmpl t = make_template(J, S, r, j, b, e);
                    uint32_t* o = s_out + tid * rpt;
                    if (fits) write_read(J, S, lv, t.x, t.src, r, j, c2, 0u, slots, BLOCK, o, stride);
                    else write_read(J, S, gv, t.x, t.src, r, j, c2, 0u, slots, BLOCK, o, stride);
                    if (J.pe) {
                        if (fits) write_read(J, S, lv, t.x2, t.src, r, j, c2, 1u, slots, BLOCK, o + 1, stride);
                        else write_read(J, S, gv, t.x2, t.src, r, j, c2, 1u, slots, BLOCK, o + 1, stride);
                    }
                }
                __syncthreads();
                // coalesced copy: reads [first, first + n) are contiguous in the output
                const uint32_t n_t = min((uint32_t)BLOCK, (uint32_t)nj - r0);
                const uint64_t first = S.read_base + (before + r0) * rpt;
                const uint32_t total = n_t * rpt * RW;
                uint32_t* dst = L.reads_out + first * RW;
                for (uint32_t q = tid; q < total; q += BLOCK) {
                    const uint32_t rd = q / RW, wd = q - rd * RW;
                    dst[q] = s_out[wd * stride + rd];
                }
                __syncthreads();
            }
            if (fits) __syncthreads();
        }
    }
}

// ------------------------------------------------------------- host side

struct prepared {
    std::vector<job_dev> jobs;
    std::vector<seg_dev> segs;
    std::vector<item_dev> items;
    std::vector<pres_dev> pres;
    uint64_t n_reads = 0, pres_total = 0;
    scratch_builder sb;
    size_t o_jobs = 0, o_segs = 0, o_items = 0, o_pres = 0;
};

struct job_in {
    const psx_genome* g;
    const psx_forest* f;
    const psx_mutations* m;
    const psx_seq_params* p;
    uint64_t seed;
    int32_t* alt;
    int32_t* depth;
    uint8_t* present;
};

uint32_t out_samples(const psx_forest* f, const psx_seq_params* p) {
    return (p->n_samples ? p->n_samples : f->n_samples) + (p->with_normal_sample ? 1u : 0u);
}

// Validates one job and appends its descriptors; the ordering of output
// samples and chromosomes follows the oracle.
void prepare_job(psx_ctx* ctx, prepared& P, const job_in& in, uint32_t spi, std::vector<size_t>& table_offs) {
    const psx_seq_params* p = in.p;
    const psx_forest* f = in.f;
    psx_mutations* m = const_cast<psx_mutations*>(in.m);
    const psx_genome* g = in.g;
    if (!(p->coverage >= 0.0) || !std::isfinite(p->coverage)) psx_throw(PSX_EINVAL, "invalid coverage");
    if (p->read_size < 1 || p->read_size > MAX_READ) psx_throw(PSX_EINVAL, "the read size must be in [1, %u]", MAX_READ);
    if (!(p->purity >= 0.0 && p->purity <= 1.0)) psx_throw(PSX_EINVAL, "The purity must be in [0,1]");
    if (!(p->error_rate >= 0.0 && p->error_rate < 1.0))
        psx_throw(PSX_EINVAL, "The parameter \"error_rate\" must be a positive real number.");
    if (m->n_leaves != f->n_leaves) psx_throw(PSX_EINVAL, "the mutation table belongs to another forest");
    const uint32_t ploidy = m->ploidy;
    if ((uint64_t)f->n_leaves * ploidy >= (uint64_t(1) << 32)) psx_throw(PSX_EINVAL, "too many cell copies");
    const uint32_t L = p->read_size;
    const bool pe = p->insert_size_mean > 0;
    psx_itable ins, err;
    uint32_t reach = L;
    if (pe) {
        ins = psx_insert_table(p->insert_size_mean, p->insert_size_stddev);
        const uint32_t imax = ins.k0 + (uint32_t)ins.T.size() - 1;
        reach = std::max(reach, imax);
        if (reach > MAX_READ) psx_throw(PSX_EINVAL, "insert sizes above %u are not supported", MAX_READ);
    }
    if (p->error_rate > 0.0) err = psx_error_table(L, p->error_rate);
    std::vector<uint8_t> sel(g->n_chr, p->n_chromosomes == 0 ? 1 : 0);
    for (uint32_t k = 0; k < p->n_chromosomes; ++k) {
        if (p->chromosomes[k] >= g->n_chr) psx_throw(PSX_EINVAL, "unknown chromosome %u", p->chromosomes[k]);
        sel[p->chromosomes[k]] = 1;
    }
    const uint32_t n_req = p->n_samples ? p->n_samples : f->n_samples;
    std::vector<uint32_t> req(n_req);
    for (uint32_t k = 0; k < n_req; ++k) {
        req[k] = p->n_samples ? p->samples[k] : k;
        if (req[k] >= f->n_samples) psx_throw(PSX_EINVAL, "unknown sample %u", req[k]);
        if (f->sample_begin[req[k] + 1] == f->sample_begin[req[k]]) psx_throw(PSX_EINVAL, "sample %u has no cells", req[k]);
    }
    if (m->chr_len != g->len) psx_mutations_index(ctx, m, g->len);
    const uint32_t n_out = n_req + (p->with_normal_sample ? 1u : 0u);
    const uint32_t job = (uint32_t)P.jobs.size();
    job_dev J{};
    J.seed = in.seed;
    J.pos = m->d_pos.as<uint32_t>(); J.lo = m->d_lo.as<uint32_t>(); J.hi = m->d_hi.as<uint32_t>();
    J.info = m->d_info.as<uint32_t>(); J.bucket = m->d_bucket.as<uint32_t>();
    J.sample_leaves = f->d_sample_leaves.as<uint32_t>();
    J.alt = in.alt; J.depth = in.depth; J.present = in.present;
    J.packed = g->packed.as<uint32_t>();
    J.n_mut = m->n;
    J.thr_normal = (uint64_t)((1.0 - p->purity) * 4294967296.0);
    J.ins_k0 = ins.k0; J.ins_n = (uint32_t)ins.T.size();
    J.err_n = (uint32_t)err.T.size();
    J.L = L; J.pe = pe ? 1u : 0u; J.ploidy = ploidy; J.n_leaves = f->n_leaves;
    J.preneo = p->preneoplastic_in_normal ? 1u : 0u; J.reach = reach; J.n_out = n_out;
    J.purity_lt1 = p->purity < 1.0 ? 1u : 0u;
    table_offs.push_back(ins.T.empty() ? SIZE_MAX : P.sb.put_vec(ins.T));
    table_offs.push_back(err.T.empty() ? SIZE_MAX : P.sb.put_vec(err.T));
    P.jobs.push_back(J);
    for (uint32_t so = 0; so < n_out; ++so) {
        const bool normal = so == n_req;
        const uint32_t leaf_off = normal ? 0u : f->sample_begin[req[so]];
        const uint32_t n_cells = normal ? 0u : f->sample_begin[req[so] + 1] - f->sample_begin[req[so]];
        if (m->n) {
            P.pres.push_back(pres_dev{P.pres_total, job, so, leaf_off, n_cells});
            P.pres_total += m->n;
        }
        for (uint32_t c = 0; c < g->n_chr; ++c) {
            if (!sel[c] || g->len[c] < L) continue;
            seg_dev S{};
            S.T = (uint64_t)std::llround(p->coverage * (double)g->len[c] / (double)(pe ? 2 * L : L));
            S.out_off = (uint64_t)so * m->n;
            S.read_base = P.n_reads;
            S.job = job; S.so = so; S.chr = c; S.len = g->len[c];
            S.W = g->len[c] - L + 1;
            S.J = (S.W + PSX_STRATUM - 1) / PSX_STRATUM;
            S.bucket_off = (uint32_t)m->bucket_off[c];
            S.n_bucket = (uint32_t)(m->bucket_off[c + 1] - m->bucket_off[c]);
            S.leaf_off = leaf_off; S.n_cells = n_cells;
            S.chr_woff = g->woff[c];
            P.n_reads += S.T * (pe ? 2u : 1u);
            const uint32_t seg = (uint32_t)P.segs.size();
            P.segs.push_back(S);
            if (S.T == 0) continue;
            for (uint32_t j0 = 0; j0 < S.J; j0 += spi) P.items.push_back(item_dev{seg, j0, std::min(S.J, j0 + spi), 0});
        }
    }
}

void finalize(psx_ctx* ctx, prepared& P, const std::vector<size_t>& table_offs) {
    // descriptors: tables first (already in sb), then jobs with patched pointers
    P.o_segs = P.sb.put_vec(P.segs);
    P.o_items = P.sb.put_vec(P.items);
    P.o_pres = P.sb.put_vec(P.pres);
    P.o_jobs = P.sb.put_vec(P.jobs);
    ctx->scratch.reserve(P.sb.host.size() + 256);
    uint8_t* base = (uint8_t*)ctx->scratch.p;
    job_dev* hj = (job_dev*)(P.sb.host.data() + P.o_jobs);
    for (size_t k = 0; k < P.jobs.size(); ++k) {
        const size_t oi = table_offs[2 * k], oe = table_offs[2 * k + 1];
        hj[k].ins_T = oi == SIZE_MAX ? nullptr : (const uint64_t*)(base + oi);
        hj[k].err_T = oe == SIZE_MAX ? nullptr : (const uint64_t*)(base + oe);
    }
    psx_scratch_upload(ctx, P.sb);
}

template <int BLOCK, int CAP>
void launch_count_t(psx_ctx* ctx, const seq_launch& L, uint32_t grid) {
    hipLaunchKernelGGL((k_seq_count<BLOCK, CAP>), dim3(grid), dim3(BLOCK), 0, ctx->stream, L);
}

template <int BLOCK, int CAP>
void launch_reads_t(psx_ctx* ctx, const seq_launch& L, uint32_t grid, size_t dyn) {
    hipLaunchKernelGGL((k_seq_reads<BLOCK, CAP>), dim3(grid), dim3(BLOCK), dyn, ctx->stream, L);
}

psx_launch_config default_config() { return psx_launch_config{256, 1024, 1, 8, 0}; }

void check_config(const psx_launch_config& c) {
    const bool block_ok = c.block_threads == 256 || c.block_threads == 512;
    const bool cap_ok = c.lds_records == 512 || c.lds_records == 1024 || c.lds_records == 2048;
    if (!block_ok || !cap_ok || c.strata_per_item < 1 || c.blocks_per_cu < 1 || c.blocks_per_cu > 64)
        psx_throw(PSX_EINVAL, "unsupported launch configuration (block %u, lds %u, strata %u, blocks/CU %u)",
                  c.block_threads, c.lds_records, c.strata_per_item, c.blocks_per_cu);
}

void launch_count(psx_ctx* ctx, const psx_launch_config& c, const seq_launch& L, uint32_t grid) {
#define PSX_CASE(B, C) \
    if (c.block_threads == B && c.lds_records == C) return launch_count_t<B, C>(ctx, L, grid);
    PSX_CASE(256, 512) PSX_CASE(256, 1024) PSX_CASE(256, 2048)
    PSX_CASE(512, 512) PSX_CASE(512, 1024) PSX_CASE(512, 2048)
#undef PSX_CASE
    psx_throw(PSX_EINVAL, "unsupported launch configuration");
}

struct run_result {
    uint64_t n_reads;
    double kernel_ms;
};

// Uploads descriptors, runs init + count kernels (optionally via a hipGraph)
// and times the count kernel with HIP events on the context stream.
run_result run_count(psx_ctx* ctx, const std::vector<job_in>& jobs, const psx_launch_config& cfg) {
    check_config(cfg);
    PSX_HIP(hipSetDevice(ctx->device));
    prepared P;
    std::vector<size_t> table_offs;
    for (const auto& j : jobs) prepare_job(ctx, P, j, cfg.strata_per_item, table_offs);
    finalize(ctx, P, table_offs);
    uint8_t* base = (uint8_t*)ctx->scratch.p;
    ctx->queue.reserve(256);
    seq_launch L{};
    L.jobs = (const job_dev*)(base + P.o_jobs);
    L.segs = (const seg_dev*)(base + P.o_segs);
    L.items = (const item_dev*)(base + P.o_items);
    L.n_items = P.items.size();
    L.queue = ctx->queue.as<unsigned long long>();
    const uint64_t want = (uint64_t)ctx->n_cu * cfg.blocks_per_cu;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, P.items.size()));
    const uint32_t init_grid = (uint32_t)std::max<uint64_t>(1, (P.pres_total + 255) / 256);
    auto enqueue = [&] {
        hipLaunchKernelGGL(k_seq_init, dim3(init_grid), dim3(256), 0, ctx->stream, L.jobs,
                           (const pres_dev*)(base + P.o_pres), (uint32_t)P.pres.size(), P.pres_total, L.queue);
        PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        launch_count(ctx, cfg, L, grid);
        PSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    };
    if (cfg.use_graph) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        PSX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipLaunchKernelGGL(k_seq_init, dim3(init_grid), dim3(256), 0, ctx->stream, L.jobs,
                           (const pres_dev*)(base + P.o_pres), (uint32_t)P.pres.size(), P.pres_total, L.queue);
        launch_count(ctx, cfg, L, grid);
        PSX_HIP(hipStreamEndCapture(ctx->stream, &graph));
        PSX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        PSX_HIP(hipGraphLaunch(exec, ctx->stream));
        PSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
    } else {
        enqueue();
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    }
    float ms = 0.f;
    PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    return run_result{P.n_reads, (double)ms};
}

struct dev_out {
    dbuf alt, depth, present;
};

}  // namespace

extern "C" uint32_t psx_seq_num_samples(const psx_forest* f, const psx_seq_params* p) {
    if (!f || !p) return 0;
    return out_samples(f, p);
}

extern "C" psx_status psx_seq_num_reads(const psx_genome* g, const psx_forest* f, const psx_seq_params* p,
                                        uint64_t* n_reads) {
    if (!g || !f || !p || !n_reads) return PSX_EINVAL;
    return psx_guard(nullptr, [&] {
        if (p->read_size < 1) psx_throw(PSX_EINVAL, "invalid read size");
        const bool pe = p->insert_size_mean > 0;
        const uint32_t L = p->read_size;
        std::vector<uint8_t> sel(g->n_chr, p->n_chromosomes == 0 ? 1 : 0);
        for (uint32_t k = 0; k < p->n_chromosomes; ++k)
            if (p->chromosomes[k] < g->n_chr) sel[p->chromosomes[k]] = 1;
        uint64_t total = 0;
        for (uint32_t c = 0; c < g->n_chr; ++c) {
            if (!sel[c] || g->len[c] < L) continue;
            total += (uint64_t)std::llround(p->coverage * (double)g->len[c] / (double)(pe ? 2 * L : L)) * (pe ? 2u : 1u);
        }
        *n_reads = total * out_samples(f, p);
    });
}

extern "C" psx_status psx_simulate_seq_device(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                              const psx_mutations* m, const psx_seq_params* p, uint64_t seed,
                                              int32_t* d_alt, int32_t* d_depth, uint8_t* d_present,
                                              uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p || !d_alt || !d_depth || !d_present) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        const psx_launch_config cfg = ctx->tuned ? ctx->tuned_cfg : default_config();
        psx_launch_config c = cfg;
        c.use_graph = 0;
        run_result r = run_count(ctx, {job_in{g, f, m, p, seed, d_alt, d_depth, d_present}}, c);
        if (n_reads) *n_reads = r.n_reads;
    });
}

extern "C" psx_status psx_simulate_seq(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                       const psx_mutations* m, const psx_seq_params* p, uint64_t seed, int32_t* alt,
                                       int32_t* depth, uint8_t* present, uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        const uint64_t n = (uint64_t)out_samples(f, p) * m->n;
        dev_out o;
        o.alt.alloc(n * 4);
        o.depth.alloc(n * 4);
        o.present.alloc(n);
        const psx_launch_config cfg = ctx->tuned ? ctx->tuned_cfg : default_config();
        psx_launch_config c = cfg;
        c.use_graph = 0;
        run_result r = run_count(ctx, {job_in{g, f, m, p, seed, o.alt.as<int32_t>(), o.depth.as<int32_t>(),
                                              o.present.as<uint8_t>()}}, c);
        if (n) {
            if (alt) PSX_HIP(hipMemcpyAsync(alt, o.alt.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
            if (depth) PSX_HIP(hipMemcpyAsync(depth, o.depth.p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
            if (present) PSX_HIP(hipMemcpyAsync(present, o.present.p, n, hipMemcpyDeviceToHost, ctx->stream));
        }
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        if (n_reads) *n_reads = r.n_reads;
    });
}

extern "C" uint32_t psx_read_words(const psx_seq_params* p) {
    return p ? 2u + (p->read_size + 15u) / 16u : 0u;
}

extern "C" psx_status psx_simulate_seq_reads(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                             const psx_mutations* m, const psx_seq_params* p, uint64_t seed,
                                             uint32_t* d_reads, uint64_t capacity_reads, uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p || !d_reads) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        const psx_launch_config cfg{256, 1024, 1, 4, 0};
        prepared P;
        std::vector<size_t> table_offs;
        // presence/counters are not produced here; point them at a scratch dummy
        job_in in{g, f, m, p, seed, nullptr, nullptr, nullptr};
        prepare_job(ctx, P, in, cfg.strata_per_item, table_offs);
        if (P.n_reads > capacity_reads)
            psx_throw(PSX_EINVAL, "the read buffer holds %llu reads, the run simulates %llu",
                      (unsigned long long)capacity_reads, (unsigned long long)P.n_reads);
        finalize(ctx, P, table_offs);
        uint8_t* base = (uint8_t*)ctx->scratch.p;
        ctx->queue.reserve(256);
        PSX_HIP(hipMemsetAsync(ctx->queue.p, 0, 8, ctx->stream));
        seq_launch L{};
        L.jobs = (const job_dev*)(base + P.o_jobs);
        L.segs = (const seg_dev*)(base + P.o_segs);
        L.items = (const item_dev*)(base + P.o_items);
        L.n_items = P.items.size();
        L.queue = ctx->queue.as<unsigned long long>();
        L.reads_out = d_reads;
        L.read_words = psx_read_words(p);
        const uint32_t rpt = p->insert_size_mean > 0 ? 2u : 1u;
        const size_t dyn = (size_t)cfg.block_threads * rpt * L.read_words * 4u;
        const uint64_t want = (uint64_t)ctx->n_cu * cfg.blocks_per_cu;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, P.items.size()));
        PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        if (!P.items.empty()) launch_reads_t<256, 1024>(ctx, L, grid, dyn);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_kernel_ms = ms;
        if (n_reads) *n_reads = P.n_reads;
    });
}

extern "C" psx_status psx_simulate_seq_batch(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                             const psx_launch_config* cfg, uint64_t* n_reads) {
    if (!ctx || (!jobs && n_jobs)) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        std::vector<job_in> in;
        for (uint32_t k = 0; k < n_jobs; ++k) {
            const psx_seq_job& J = jobs[k];
            if (!J.genome || !J.forest || !J.mutations || !J.d_alt || !J.d_depth || !J.d_present)
                psx_throw(PSX_EINVAL, "job %u is incomplete", k);
            in.push_back(job_in{J.genome, J.forest, J.mutations, &J.params, J.seed, J.d_alt, J.d_depth, J.d_present});
        }
        psx_launch_config c = cfg ? *cfg : (ctx->tuned ? ctx->tuned_cfg : default_config());
        if (!cfg) c.use_graph = 1;
        run_result r = in.empty() ? run_result{0, 0.0} : run_count(ctx, in, c);
        if (n_reads) *n_reads = r.n_reads;
    });
}

extern "C" psx_status psx_autotune_seq(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                       psx_launch_config* best, double* best_ms) {
    if (!ctx || !jobs || !n_jobs || !best) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        std::vector<job_in> in;
        for (uint32_t k = 0; k < n_jobs; ++k) {
            const psx_seq_job& J = jobs[k];
            in.push_back(job_in{J.genome, J.forest, J.mutations, &J.params, J.seed, J.d_alt, J.d_depth, J.d_present});
        }
        const uint32_t blocks[] = {256, 512};
        const uint32_t caps[] = {512, 1024, 2048};
        const uint32_t spis[] = {1, 4};
        const uint32_t bpcs[] = {2, 4, 8, 16};
        double best_t = 1e300;
        psx_launch_config bc = default_config();
        for (uint32_t b : blocks)
            for (uint32_t cap : caps)
                for (uint32_t spi : spis)
                    for (uint32_t bpc : bpcs) {
                        if (b == 512 && bpc > 8) continue;
                        psx_launch_config c{b, cap, spi, bpc, 0};
                        double t = 0.0;
                        for (int rep = 0; rep < 2; ++rep) t += run_count(ctx, in, c).kernel_ms;
                        if (t < best_t) { best_t = t; bc = c; }
                    }
        ctx->tuned = true;
        ctx->tuned_cfg = bc;
        *best = bc;
        if (best_ms) *best_ms = best_t / 2.0;
    });
}
