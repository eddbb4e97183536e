// Device side of simulate_seq on gfx950: descriptors, the count-only kernel
// and the read-materialising kernel. Included once, by psx_seq.hip.
//
// Work decomposition. A segment is one (job, output sample, chromosome); its
// W = len - L + 1 template starts are cut into strata of PSX_STRATUM starts,
// and stratum j owns a fixed number of templates (part of the sampling model,
// so results do not depend on the launch). A work item is a run of strata of
// one segment; persistent blocks pull items from a device counter.
//
// Per stratum a block stages in LDS the mutation records the stratum's reads
// can reach (pos, leaf range, packed info) with a 32-base bucket index, plus
// the job's inversion tables; reads are generated from counters and looked up
// in LDS, counts go to LDS counters and the block flushes them with one global
// atomic per touched record. All counting is integer.
#pragma once
#include "psx_cna.h"
#include "psx_internal.h"

namespace psxk {

constexpr uint32_t NB_SHIFT = 5;      // 32-base LDS buckets
constexpr uint32_t MAX_READ = 4096;   // read size and insert reach limit
constexpr uint32_t NB_MAX = ((PSX_STRATUM + MAX_READ) >> NB_SHIFT) + 2;
constexpr uint32_t MAX_INS = 640;     // compact insert-size thresholds in LDS
constexpr uint32_t MAX_ERRT = PSX_ERR_KCAP + 1;
constexpr uint32_t REF_MAX = (PSX_STRATUM + MAX_READ) / 16 + 4;
constexpr uint32_t UNRESOLVED = 0xFFFFFFFFu;

// Descriptors are read-only for the whole launch: reading them through the
// constant address space lets the compiler use scalar loads, so the fields
// live in SGPRs instead of taking VGPRs from every lane.
template <class T>
__device__ __forceinline__ T uniform_at(const T* p, uint64_t i) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(4))) T*)p)[i];
#else
    return p[i];
#endif
}

struct job_dev {
    uint64_t seed;
    const uint32_t* pos;
    const uint32_t* lo;
    const uint32_t* hi;
    const uint32_t* info;
    const uint32_t* bucket;
    const uint32_t* sample_leaves;
    const uint4* leaf_count;     // per sample and 64 leaf ranks: {leaves below, membership bits, 0}
    int32_t* alt;
    int32_t* depth;
    uint8_t* present;
    const uint32_t* ins_T;       // compact insert-size thresholds
    const uint16_t* ins_guide;
    const uint32_t* err_T;       // compact thresholds of the errors per template
    const uint32_t* packed;      // reference (read materialisation)
    uint64_t n_mut;
    uint64_t thr_normal;         // floor((1 - purity) 2^32)
    uint32_t ins_k0, ins_n;
    uint32_t err_n, err_k0;      // error table (err_n == 0: no sequencing errors)
    uint32_t err_N;              // bases per template: (pe + 1) L
    uint32_t inv_L;              // ceil(2^32 / L): b / L = mulhi(b, inv_L) for b < 2^20
    uint32_t L, pe, ploidy, n_leaves, preneo, reach;
    uint32_t n_out;
    uint32_t purity_lt1;
    uint32_t leaf_blocks;        // leaf_count entries per sample
    const psx_cna_dev* cna;      // copy-number events of every chromosome, per segment ranges
    unsigned long long* acc;     // reads of accepted templates (jobs with copy-number events)
    const uint64_t* cn_row;      // read materialisation: first read of each copy-number stratum (seg st_off + j)
};

struct seg_dev {
    uint64_t T;             // templates of the segment
    uint64_t out_off;       // so * n_mut
    uint64_t read_base;     // first read index of the segment (materialisation)
    uint64_t chr_woff;      // first packed word of the chromosome
    uint32_t job, so, chr, len;
    uint32_t W, J, bucket_off, n_bucket;
    uint32_t leaf_off, n_cells;  // n_cells == 0: normal sample
    uint32_t cna_off, n_cna;     // the chromosome's copy-number events (J.cna + cna_off)
    uint32_t kmax;               // copies drawn per template: the largest copy number, >= ploidy
    uint32_t first_item;         // its work items: first_item .. first_item + ceil(J / strata per item) - 1
    uint32_t st_off;             // segments with copy-number events: index of its first stratum in J.cn_row
};

struct item_dev {
    uint32_t seg, j0, j1, pad;
};

// The work items of a batch, written on the device from the segments (the
// host lists segments only): item q belongs to the last segment whose first
// item is at or below q (a segment without templates has no items and shares
// its successor's first item) and covers spi strata from its first.
__global__ void __launch_bounds__(256) k_seq_items(const seg_dev* __restrict__ segs, uint32_t n_segs,
                                                   item_dev* __restrict__ items, uint32_t n_items, uint32_t spi) {
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= n_items) return;
    uint32_t a = 0, z = n_segs;
    while (a < z) {
        const uint32_t mid = (a + z) >> 1;
        if (segs[mid].first_item <= q) a = mid + 1; else z = mid;
    }
    const seg_dev& S = segs[a - 1];
    const uint32_t j0 = (q - S.first_item) * spi;
    items[q] = item_dev{a - 1, j0, min(S.J, j0 + spi), 0u};
}

struct pres_dev {  // one per (job, output sample) for the init kernel
    uint64_t begin;        // global index of the first (so, m) entry
    uint32_t job, so, leaf_off, n_cells;
    uint32_t sample, pad;  // the forest sample of the output sample (tumour samples)
};

struct seq_launch {
    const job_dev* jobs;
    const seg_dev* segs;
    const item_dev* items;
    uint64_t n_items;
    unsigned long long* queue;
    uint32_t* reads_out;     // materialisation only
    uint32_t read_words;
    uint32_t out_vec4;       // read_words % 4 == 0 and reads_out 16-byte aligned
};

// The job's inversion tables in LDS; the insert-size guide keeps 2^GB of the
// host table's PSX_GUIDE entries (every 2^(PSX_GUIDE_BITS - GB)-th, the same
// guide at a coarser step), so a kernel short of LDS can hold a smaller one.
template <uint32_t GB>
struct lds_tables_t {
    static constexpr uint32_t GUIDE_BITS = GB;
    uint32_t ins[MAX_INS];
    uint16_t guide[1u << GB];
    uint32_t err[MAX_ERRT];
};
using lds_tables = lds_tables_t<PSX_GUIDE_BITS>;

template <class TAB>
__device__ __forceinline__ void load_tables(const job_dev& J, TAB& t, uint32_t tid, uint32_t block) {
    constexpr uint32_t GB = TAB::GUIDE_BITS;
    static_assert(GB <= PSX_GUIDE_BITS, "the host guide is the finest");
    if (J.pe) {
        for (uint32_t i = tid; i < J.ins_n; i += block) t.ins[i] = J.ins_T[i];
        for (uint32_t i = tid; i < (1u << GB); i += block) t.guide[i] = J.ins_guide[i << (PSX_GUIDE_BITS - GB)];
    }
    for (uint32_t i = tid; i < J.err_n; i += block) t.err[i] = J.err_T[i];
    // the error-count pick reads thresholds 0..3 whatever the table's length
    for (uint32_t i = J.err_n + tid; i < 4u; i += block) t.err[i] = 0xFFFFFFFFu;
}

// LDS writes of a wave visible to its other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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
    // a record index at or before the first record at or after x
    __device__ __forceinline__ uint32_t start(uint32_t x) const { return bucket[(x - b) >> NB_SHIFT]; }
    __device__ __forceinline__ uint32_t at(uint32_t i) const { return min(i, n - 1u); }  // n > 0
    __device__ __forceinline__ void add_dep(uint32_t q) const { atomicAdd(&dep[q], 1u); }
    __device__ __forceinline__ void add_alt(uint32_t q) const { atomicAdd(&alt[q], 1u); }
    __device__ __forceinline__ void add_alt_n(uint32_t q, int d) const { atomicAdd(&alt[q], (uint32_t)d); }
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
    __device__ __forceinline__ uint32_t first_u(uint32_t x) const { return first(x); }
    __device__ __forceinline__ uint32_t start(uint32_t x) const { return first(x); }
    __device__ __forceinline__ uint32_t at(uint32_t i) const { return min(i, n - 1u); }  // n > 0
    __device__ __forceinline__ void add_dep(uint32_t q) const { atomicAdd(&dep[q], 1); }
    __device__ __forceinline__ void add_alt(uint32_t q) const { atomicAdd(&alt[q], 1); }
    __device__ __forceinline__ void add_alt_n(uint32_t q, int d) const { atomicAdd(&alt[q], d); }
};

// Template: start of mate 1 and mate 2, source cell and its leaf, and copy;
// ok == 0 when a copy-number draw rejects it.
struct tmpl {
    uint32_t x, x2, cell, allele, leaf, ok;
};

// Copy numbers of a leaf at base x of the segment's chromosome, from its
// events read with scalar loads (the list is the same for every lane);
// returns the total. Oracle: copies.
__device__ __forceinline__ uint32_t cna_copies(const psx_cna_dev* ev, uint32_t n, uint32_t leaf, uint32_t x,
                                               uint32_t ploidy, uint32_t* cn) {
#pragma unroll
    for (uint32_t a = 0; a < PSX_MAX_CNA_PLOIDY; ++a) cn[a] = a < ploidy ? 1u : 0u;
    for (uint32_t i = 0; i < n; ++i) {
        const psx_cna_dev e = uniform_at(ev, i);
        const bool hit = leaf >= e.lo && leaf < e.hi && x >= e.start && x < e.end;
        const uint32_t a = e.code & 0xFFu, kind = (e.code >> 8) & 0xFFu, c = e.code >> 16;
#pragma unroll
        for (uint32_t q = 0; q < PSX_MAX_CNA_PLOIDY; ++q) {
            const uint32_t v = cn[q];
            const uint32_t nv = kind == PSX_CNA_WGD ? 2u * v : (q != a ? v : (kind == PSX_CNA_AMP ? v + c : 0u));
            cn[q] = hit ? nv : v;
        }
    }
    uint32_t total = 0;
#pragma unroll
    for (uint32_t a = 0; a < PSX_MAX_CNA_PLOIDY; ++a) total += cn[a];
    return total;
}

// cna_copies for ploidy 2: copies of lineage 0 | lineage 1 << 16 (copy
// numbers stay below 2^16: the host bounds them by the largest copy number).
__device__ __forceinline__ uint32_t cna_copies2(const psx_cna_dev* ev, uint32_t n, uint32_t leaf, uint32_t x) {
    uint32_t c0 = 1u, c1 = 1u;
    for (uint32_t i = 0; i < n; ++i) {
        const psx_cna_dev e = uniform_at(ev, i);
        const bool hit = leaf >= e.lo && leaf < e.hi && x >= e.start && x < e.end;
        const uint32_t a = e.code & 0xFFu, kind = (e.code >> 8) & 0xFFu, c = e.code >> 16;
        uint32_t n0, n1;
        if (kind == PSX_CNA_WGD) {  // uniform branches: the event is the same for every lane
            n0 = 2u * c0;
            n1 = 2u * c1;
        } else if (kind == PSX_CNA_AMP) {
            n0 = a == 0u ? c0 + c : c0;
            n1 = a == 1u ? c1 + c : c1;
        } else {
            n0 = a == 0u ? 0u : c0;
            n1 = a == 1u ? 0u : c1;
        }
        c0 = hit ? n0 : c0;
        c1 = hit ? n1 : c1;
    }
    return c0 | (c1 << 16);
}

__device__ __forceinline__ uint32_t leaf_of(const job_dev& J, const seg_dev& S, tmpl& t) {
    if (t.leaf == UNRESOLVED) t.leaf = J.sample_leaves[S.leaf_off + t.cell];
    return t.leaf;
}

__device__ __forceinline__ bool carried(const job_dev& J, uint32_t info, uint32_t lo, uint32_t hi, uint32_t leaf) {
    if (leaf == J.n_leaves) {
        const uint32_t nat = PSX_INFO_NATURE(info);
        return nat == PSX_GERMINAL || (nat == PSX_PRENEOPLASTIC && J.preneo);
    }
    return lo <= leaf && leaf < hi;
}

// psx_guided_pick with a wave-uniform binary search inside the guide bucket:
// the answer lies in [guide[g], guide[g + 1]] (the last bucket ends at n - 1),
// and the tail buckets of a peaked distribution hold tens of entries, which a
// linear scan would walk for the whole wave. Same result as psx_guided_pick.
template <uint32_t GB>
__device__ __forceinline__ uint32_t guided_pick_u(const uint32_t* T32, const uint16_t* guide, uint32_t n, uint32_t u) {
    const uint32_t g = u >> (32u - GB);
    uint32_t lo = guide[g];
    uint32_t hi = g < (1u << GB) - 1u ? guide[g + 1u] : n - 1u;
    while (__ballot(lo < hi)) {
        const uint32_t mid = (lo + hi) >> 1;
        const bool below = u < T32[mid];
        const bool on = lo < hi;
        hi = on && below ? mid : hi;
        lo = on && !below ? mid + 1u : lo;
    }
    return lo;
}

// Template r of stratum j without its mate (oracle: template_source): start,
// source and, with copy-number events, whether the template is kept. w gets
// the template draw (its word w is the insert size).
template <bool CNA = false>
__device__ __forceinline__ tmpl template_source(const job_dev& J, const seg_dev& S, uint32_t r, uint32_t j,
                                                uint32_t b, uint32_t e, psx_u4& w) {
    const uint32_t c2 = (S.chr << 16) | S.so;
    w = psx_draw_u(J.seed, r, j, c2, PSX_P_TEMPLATE << 24);
    tmpl t;
    t.x = b + psx_uniform32(w.x, e - b);
    t.ok = 1u;
    if (!CNA || S.n_cna == 0) {
        if (S.n_cells == 0 || (uint64_t)w.y < J.thr_normal) {
            t.leaf = J.n_leaves;
            t.cell = 0;
            t.allele = psx_uniform32(w.z, J.ploidy);
        } else {
            const uint32_t X = psx_uniform32(w.z, S.n_cells * J.ploidy);
            if (J.ploidy == 2) { t.cell = X >> 1; t.allele = X & 1u; }
            else { t.cell = X / J.ploidy; t.allele = X - t.cell * J.ploidy; }
            t.leaf = J.sample_leaves[S.leaf_off + t.cell];  // issued early: its latency overlaps the draws
        }
    } else {
        // one of kmax copy slots per cell; the template is kept when the slot
        // is one of the source's copies at x, whose lineage it then reads
        const uint32_t K = S.kmax;
        if (S.n_cells == 0 || (uint64_t)w.y < J.thr_normal) {
            const uint32_t k = psx_uniform32(w.z, K);
            t.leaf = J.n_leaves;
            t.cell = 0;
            t.allele = k;
            t.ok = k < J.ploidy;
        } else {
            const uint32_t X = psx_uniform32(w.z, S.n_cells * K);
            t.cell = X / K;
            const uint32_t k = X - t.cell * K;
            t.leaf = J.sample_leaves[S.leaf_off + t.cell];
            if (J.ploidy == 2u) {
                // diploid: the two lineages' copy numbers in two registers
                const uint32_t C0 = cna_copies2(J.cna + S.cna_off, S.n_cna, t.leaf, t.x);
                const uint32_t c0 = C0 & 0xFFFFu, c1 = C0 >> 16;
                t.allele = k >= c0 ? 1u : 0u;
                t.ok = k < c0 + c1;
            } else {
                uint32_t cn[PSX_MAX_CNA_PLOIDY];
                const uint32_t C = cna_copies(J.cna + S.cna_off, S.n_cna, t.leaf, t.x, J.ploidy, cn);
                uint32_t cum = 0, al = 0;
#pragma unroll
                for (uint32_t a = 0; a < PSX_MAX_CNA_PLOIDY; ++a) {
                    cum += cn[a];
                    al += (a + 1 < J.ploidy && k >= cum) ? 1u : 0u;
                }
                t.allele = al;
                t.ok = k < C;
            }
        }
    }
    t.x2 = t.x;
    return t;
}

// Template r of stratum j (oracle: the template loop of orc_simulate_seq).
template <bool CNA = false, class TAB>
__device__ __forceinline__ tmpl make_template(const job_dev& J, const seg_dev& S, const TAB& tab, uint32_t r,
                                              uint32_t j, uint32_t b, uint32_t e) {
    psx_u4 w;
    tmpl t = template_source<CNA>(J, S, r, j, b, e, w);
    if (J.pe) {
        uint32_t I = J.ins_k0 + guided_pick_u<TAB::GUIDE_BITS>(tab.ins, tab.guide, J.ins_n, w.w);
        if (I < J.L) I = J.L;
        uint64_t x2 = (uint64_t)t.x + I - J.L;
        if (x2 > S.len - J.L) x2 = S.len - J.L;
        t.x2 = (uint32_t)x2;
    }
    return t;
}

// Sequencing errors of a template (oracle: template_errors). Most templates
// are settled by their first draw w_0: K <= 3 and the stream words x, z, w of
// w_0 give distinct base indices; their errors travel in a two-word summary,
// each error as e = b | sub << 13 (b < 2 L <= 8192 the template's base
// index, sub the substitution index). The others (slow) are generated by the
// whole wave (wave_template_errors).
struct terr {
    uint32_t w01;  // e_0 | e_1 << 16
    uint32_t w2n;  // e_2 | n << 16 | slow << 18
    __device__ __forceinline__ uint32_t n() const { return (w2n >> 16) & 3u; }
    __device__ __forceinline__ bool slow() const { return (w2n >> 18) & 1u; }
};

// The summary of template r of stratum j; lanes with want == false get none.
// Every lane steps through the same count loop.
template <class TAB>
__device__ __forceinline__ terr template_errors(const job_dev& J, const TAB& tab, uint32_t r, uint32_t j,
                                                uint32_t c2, bool want) {
    const psx_u4 w = psx_draw_u(J.seed, r, j, c2, PSX_P_ERROR << 24);
    // K = err_k0 + the number of thresholds k < err_n - 1 with u >= T[k] (T is
    // non-decreasing, every kept threshold but the last below 2^32, the last and
    // the padding 2^32 - 1); only K <= 3 matters here, so four compares against
    // the first thresholds (read by every lane alike) give min(K, 4) without a
    // loop, and capping at err_n - 1 drops the last threshold and the padding,
    // which only u = 2^32 - 1 reaches
    uint32_t k = 0;
#pragma unroll
    for (uint32_t q = 0; q < 4u; ++q) k += w.y >= tab.err[q] ? 1u : 0u;
    k = min(k, J.err_n ? J.err_n - 1u : 0u);
    const uint32_t K = want ? J.err_k0 + k : 0u;
    const uint32_t b0 = psx_mulhi32(w.x, J.err_N), b1 = psx_mulhi32(w.z, J.err_N), b2 = psx_mulhi32(w.w, J.err_N);
    const bool slow = K > 3u || (K >= 2u && b1 == b0) || (K >= 3u && (b2 == b0 || b2 == b1));
    const uint32_t n = slow ? 0u : K;
    const uint32_t e0 = b0 | ((((w.x & 0xFFFFu) * 3u) >> 16) << 13);
    const uint32_t e1 = b1 | ((((w.z & 0xFFFFu) * 3u) >> 16) << 13);
    const uint32_t e2 = b2 | ((((w.w & 0xFFFFu) * 3u) >> 16) << 13);
    terr e;
    e.w01 = (n > 0u ? e0 : 0u) | ((n > 1u ? e1 : 0u) << 16);
    e.w2n = (n > 2u ? e2 : 0u) | (n << 16) | ((slow ? 1u : 0u) << 18);
    return e;
}

__device__ __forceinline__ terr shfl_terr(const terr& e, uint32_t src) {
    return terr{(uint32_t)__shfl((int)e.w01, (int)src), (uint32_t)__shfl((int)e.w2n, (int)src)};
}

// The observed base at template base index b whose true base is tb, with the
// summary's errors (none for a slow template).
__device__ __forceinline__ uint32_t observe(const terr& e, uint32_t b, uint32_t tb) {
    const uint32_t n = e.n();
    const uint32_t e0 = e.w01 & 0xFFFFu, e1 = e.w01 >> 16, e2 = e.w2n & 0xFFFFu;
    uint32_t sub = 4u;
    sub = n > 0u && (e0 & 0x1FFFu) == b ? e0 >> 13 : sub;
    sub = n > 1u && (e1 & 0x1FFFu) == b ? e1 >> 13 : sub;
    sub = n > 2u && (e2 & 0x1FFFu) == b ? e2 >> 13 : sub;
    return sub == 4u ? tb : (sub < tb ? sub : sub + 1u);
}

// The substitution index of the summary's error at template base index b, or
// 4 when b has none.
__device__ __forceinline__ uint32_t error_sub(const terr& e, uint32_t b) {
    const uint32_t n = e.n();
    const uint32_t e0 = e.w01 & 0xFFFFu, e1 = e.w01 >> 16, e2 = e.w2n & 0xFFFFu;
    uint32_t sub = 4u;
    sub = n > 0u && (e0 & 0x1FFFu) == b ? e0 >> 13 : sub;
    sub = n > 1u && (e1 & 0x1FFFu) == b ? e1 >> 13 : sub;
    sub = n > 2u && (e2 & 0x1FFFu) == b ? e2 >> 13 : sub;
    return sub;
}

// Per-wave LDS of a slow template's errors: the list, then a 2048-bit hashed
// bitmap of the base indices drawn so far.
constexpr uint32_t ELIST = PSX_ERR_KCAP + 64u;

// All errors of template r (oracle: template_errors), generated by one wave:
// the count K from word y of w_0 (a parallel compare against the
// thresholds), then the stream words in batches, lane l drawing stream word
// t0 + l (x, z, w of w_0, then the four words of w_1, w_2, ...). A word is new
// when no earlier stream word gives the same base index (an earlier equal
// index was taken or was itself a repeat of a taken one), and the first K new
// words are the errors. Repeats are rare: the bitmap flags a batch that may
// hold one, and only such a batch runs the exact check. Writes the errors to
// list[0 .. K) as b | sub << 20 and returns K (wave-uniform); every lane
// steps through the same loops.
__device__ __forceinline__ uint32_t wave_template_errors(uint64_t seed, uint32_t err_n, uint32_t err_k0, uint32_t err_N,
                                                      const uint32_t* err_T, uint32_t r, uint32_t j, uint32_t c2,
                                                      uint32_t* list, uint32_t lane) {
    constexpr uint32_t TMAX = 4u * PSX_ERR_KCAP;  // stream words available
    uint32_t* bits = list + PSX_ERR_KCAP;
    uint32_t K = 0, cnt = 0;
    for (uint32_t t0 = 0; t0 < TMAX;) {
        const uint32_t t = t0 + lane;
        const uint32_t call = t < 3u ? 0u : 1u + ((t - 3u) >> 2);
        const psx_u4 w = psx_draw_u(seed, r, j, c2, (PSX_P_ERROR << 24) | call);
        const uint32_t word = psx_word(w, t < 3u ? (t == 0u ? 0 : (int)t + 1) : (int)((t - 3u) & 3u));
        if (t0 == 0) {
            // K - err_k0 = the number of thresholds k < err_n - 1 with u >= T[k]
            const uint32_t u = (uint32_t)__builtin_amdgcn_readfirstlane((int)w.y);  // lane 0 drew w_0
            uint32_t k = 0;
            for (uint32_t i0 = 0; i0 + 1u < err_n; i0 += 64u) {
                const uint32_t i = i0 + lane;
                k += (uint32_t)__popcll(__ballot(i + 1u < err_n && u >= err_T[min(i, err_n - 1u)]));
            }
            K = err_k0 + k;
            if (K == 0) return 0;
            bits[lane] = 0u;
            wave_sync();
        }
        const uint32_t B = min(64u, K - cnt);  // lanes [0, B) may be needed
        const uint32_t x = psx_mulhi32(word, err_N);
        const bool live = lane < B && t < TMAX;
        const uint32_t h = (x * 0x9E3779B1u) >> 21;
        const uint32_t seen = live ? atomicOr(&bits[h >> 5], 1u << (h & 31u)) & (1u << (h & 31u)) : 0u;
        unsigned long long dupm = 0;
        if (__ballot(seen != 0u)) {  // exact check against the batch and the earlier batches
            for (uint32_t q = 0; q + 1u < B; ++q) {
                const uint32_t xq = (uint32_t)__builtin_amdgcn_readlane((int)x, (int)q);
                dupm |= __ballot(live && x == xq) & (~0ull << (q + 1u));
            }
            bool d = false;
            for (uint32_t q = 0; q < cnt; ++q) d |= (list[q] & 0xFFFFFu) == x;
            dupm |= __ballot(live && d);
        }
        const bool fresh = live && !((dupm >> lane) & 1ull);
        const unsigned long long m = __ballot(fresh);
        const uint32_t rank =
            cnt + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (fresh && rank < K) list[rank] = x | ((((word & 0xFFFFu) * 3u) >> 16) << 20);
        cnt = min(K, cnt + (uint32_t)__popcll(m));
        if (cnt == K) break;
        t0 += B;
        wave_sync();  // list entries of this batch seen by every lane
    }
    return cnt;
}

// Records of one position shared by several records, from index i (the
// oracle's seq_read step): the first carried SNV sets the true base, the
// read's error there (summary e; base index p + boff) substitutes it, every
// record gains depth and each SID head an alternative read when its
// alternative is observed (SNV) or carried (indel). Returns the index past
// the position.
template <class V>
__device__ __forceinline__ uint32_t count_group(const job_dev& J, const V& v, uint32_t i, uint32_t allele,
                                                uint32_t leaf, const terr& e, uint32_t boff) {
    const uint32_t p = v.pos[i];
    const uint32_t g0 = i;
    uint32_t g1 = i + 1;
    while (g1 < v.n && v.pos[g1] == p) ++g1;
    uint32_t tb = PSX_INFO_REF(v.info[g0]);
    for (uint32_t q = g0; q < g1; ++q) {
        const uint32_t in = v.info[q];
        if (PSX_INFO_KIND(in) == PSX_SNV && PSX_INFO_ALLELE(in) == allele && carried(J, in, v.lo[q], v.hi[q], leaf)) {
            tb = PSX_INFO_ALT(in);
            break;
        }
    }
    const uint32_t ob = observe(e, p + boff, tb);
    for (uint32_t q = g0; q < g1; ++q) {
        v.add_dep(q);
        const uint32_t in = v.info[q];
        if (!PSX_INFO_HEAD(in)) continue;
        if (PSX_INFO_KIND(in) == PSX_SNV) {
            if (ob == PSX_INFO_ALT(in)) v.add_alt(q);
        } else {
            bool car = false;
            for (uint32_t q2 = q; q2 < g1 && (q2 == q || !PSX_INFO_HEAD(v.info[q2])); ++q2) {
                const uint32_t in2 = v.info[q2];
                car |= PSX_INFO_ALLELE(in2) == allele && carried(J, in2, v.lo[q2], v.hi[q2], leaf);
            }
            if (car) v.add_alt(q);
        }
    }
    return g1;
}

// A sequencing error of a slow template replacing the true base at the
// position of record i (the first of its position) by substitution index
// sub: the counts took the true base there, so each SNV head changes its
// alternative count by [observed == alt] - [true == alt]. Depth and indel
// counts do not depend on the observed base.
template <class V>
__device__ __forceinline__ void error_fix(const job_dev& J, const V& v, uint32_t i, uint32_t allele, uint32_t leaf,
                                          uint32_t sub) {
    const uint32_t p = v.pos[i];
    uint32_t g1 = i + 1;
    while (g1 < v.n && v.pos[g1] == p) ++g1;
    uint32_t tb = PSX_INFO_REF(v.info[i]);
    for (uint32_t q = i; q < g1; ++q) {
        const uint32_t in = v.info[q];
        if (PSX_INFO_KIND(in) == PSX_SNV && PSX_INFO_ALLELE(in) == allele && carried(J, in, v.lo[q], v.hi[q], leaf)) {
            tb = PSX_INFO_ALT(in);
            break;
        }
    }
    const uint32_t ob = sub < tb ? sub : sub + 1u;
    for (uint32_t q = i; q < g1; ++q) {
        const uint32_t in = v.info[q];
        if (!PSX_INFO_HEAD(in) || PSX_INFO_KIND(in) != PSX_SNV) continue;
        const int d = (ob == PSX_INFO_ALT(in) ? 1 : 0) - (tb == PSX_INFO_ALT(in) ? 1 : 0);
        if (d) v.add_alt_n(q, d);
    }
}

// The slow templates among a wave's 64 (one per lane: t, e, template index
// rr, counted when valid) were counted without errors: for each, the wave
// generates its errors and each lane takes one, correcting the counts where
// the error's base holds records. A read covering that position resolved its
// template's leaf while it was counted.
template <class V, class TAB>
__device__ __forceinline__ void count_slow_errors(const job_dev& J, const V& v, const TAB& tab, const tmpl& t,
                                                  const terr& e, bool valid, uint32_t rr, uint32_t j, uint32_t c2,
                                                  uint32_t* list, uint32_t lane) {
    unsigned long long sm = __ballot(valid && e.slow());
    while (sm) {  // wave-uniform
        const uint32_t sl = (uint32_t)__ffsll((long long)sm) - 1u;
        sm &= sm - 1ull;
        const uint32_t xs = (uint32_t)__builtin_amdgcn_readlane((int)t.x, (int)sl);
        const uint32_t x2s = (uint32_t)__builtin_amdgcn_readlane((int)t.x2, (int)sl);
        const uint32_t ls = (uint32_t)__builtin_amdgcn_readlane((int)t.leaf, (int)sl);
        const uint32_t as = (uint32_t)__builtin_amdgcn_readlane((int)t.allele, (int)sl);
        const uint32_t rs = (uint32_t)__builtin_amdgcn_readlane((int)rr, (int)sl);
        wave_sync();
        const uint32_t n = wave_template_errors(J.seed, J.err_n, J.err_k0, J.err_N, tab.err, rs, j, c2, list, lane);
        wave_sync();
        for (uint32_t k0 = 0; k0 < n; k0 += 64u) {
            const uint32_t k = k0 + lane;
            if (k < n) {
                const uint32_t w = list[k];
                const uint32_t b = w & 0xFFFFFu, sub = w >> 20;
                const uint32_t m = b >= J.L ? 1u : 0u;
                const uint32_t P = (m ? x2s : xs) + b - m * J.L;
                const uint32_t i = v.first(P);
                if (i < v.n && v.pos[i] == P) error_fix(J, v, i, as, ls, sub);
            }
        }
    }
}

// One read [x, x + L) (mate `mate` of the lane's template; summary e): depth
// for every record at a covered position, an alternative read for each SID
// head whose alternative is observed (SNV) or carried (indel). Oracle:
// seq_read. A record alone at its position (every record under the
// infinite-sites model) takes one branch-free step of a wave-uniform loop; a
// position shared by several records goes through count_group. Lanes with
// valid == false take no part. Returns whether any lane's read covers a
// record (wave-uniform).
template <class V>
__device__ __forceinline__ bool count_read(const job_dev& J, const seg_dev& S, const V& v, tmpl& t, const terr& e,
                                           bool valid, uint32_t x, uint32_t mate) {
    const uint32_t end = x + J.L;
    uint32_t i = v.first_u(x);
    bool go = valid && i < v.n && v.pos[v.at(i)] < end;
    if (!__ballot(go)) return false;
    uint32_t leaf = 0;
    if (go) leaf = leaf_of(J, S, t);
    const uint32_t boff = mate * J.L - x;  // base index of position p: p + boff
    while (__ballot(go)) {
        const uint32_t ic = go ? i : 0u;
        const uint32_t p = v.pos[ic], in = v.info[ic];
        if (__ballot(go && !PSX_INFO_SOLO(in))) {
            if (go && !PSX_INFO_SOLO(in)) i = count_group(J, v, ic, t.allele, leaf, e, boff);
        }
        const bool one = go && PSX_INFO_SOLO(in);
        const bool car = PSX_INFO_ALLELE(in) == t.allele && carried(J, in, v.lo[ic], v.hi[ic], leaf);
        const bool snv = PSX_INFO_KIND(in) == PSX_SNV;
        const uint32_t tb = snv && car ? PSX_INFO_ALT(in) : PSX_INFO_REF(in);
        const uint32_t ob = observe(e, p + boff, tb);
        if (one) {
            v.add_dep(ic);
            if (PSX_INFO_HEAD(in) && (snv ? ob == PSX_INFO_ALT(in) : car)) v.add_alt(ic);
        }
        i += one ? 1u : 0u;
        go = go && i < v.n && v.pos[v.at(i)] < end;
    }
    return true;
}

// Deferred error checks of a wave (count kernel, LDS windows): a ring of
// 128 entries of three words (template r; read base index b | allele << 13 |
// record << 21; leaf), filled by the queue drains with every covered position
// whose observed base matters (an SNV head, or several records) and checked
// 64 at a time, so one draw per lane serves 64 covered positions instead of
// one per template. head is 0 or 64 and count < 64 between drains.
struct chk_ring {
    uint32_t* e;  // 3 * 128 words, entry-major
    uint32_t head, count;  // wave-uniform
};

// Checks the n <= 64 entries at the ring's head, one per lane: the entry's
// template draws its error summary, and an error at the entry's base index
// corrects the counts there (error_fix). The entries are read first, so a
// slow template's error list can use their half of the ring. Oracle:
// seq_read with the read's errors.
template <class TAB>
__device__ __forceinline__ void check_errors(const job_dev& J, const lds_view& v, const TAB& tab, uint32_t j,
                                             uint32_t c2, chk_ring& q, uint32_t n, uint32_t lane) {
    const bool on = lane < n;
    uint32_t* half = q.e + q.head * 3u;
    const uint32_t r = on ? half[lane * 3u] : 0u;
    const uint32_t w = on ? half[lane * 3u + 1u] : 0u;
    const uint32_t leaf = on ? half[lane * 3u + 2u] : 0u;
    q.head = (q.head + 64u) & 127u;
    q.count -= n;
    const uint32_t b = w & 0x1FFFu, al = (w >> 13) & 0xFFu, ri = w >> 21;
    const terr e = template_errors(J, tab, r, j, c2, on);
    const uint32_t sub = error_sub(e, b);
    if (on && sub < 4u) error_fix(J, v, ri, al, leaf, sub);
    unsigned long long sm = __ballot(on && e.slow());
    while (sm) {  // wave-uniform, rare
        const uint32_t sl = (uint32_t)__ffsll((long long)sm) - 1u;
        sm &= sm - 1ull;
        const uint32_t rs = (uint32_t)__builtin_amdgcn_readlane((int)r, (int)sl);
        const uint32_t bs = (uint32_t)__builtin_amdgcn_readlane((int)b, (int)sl);
        wave_sync();
        const uint32_t ne = wave_template_errors(J.seed, J.err_n, J.err_k0, J.err_N, tab.err, rs, j, c2, half, lane);
        wave_sync();
        uint32_t found = 4u;  // at most one of the template's errors is at bs
        for (uint32_t k0 = 0; k0 < ne; k0 += 64u) {
            const uint32_t k = k0 + lane;
            const uint32_t le = k < ne ? half[k] : 0u;
            found = k < ne && (le & 0xFFFFFu) == bs ? le >> 20 : found;
        }
        const unsigned long long fm = __ballot(found < 4u);
        if (fm) {
            const uint32_t fs = (uint32_t)__builtin_amdgcn_readlane((int)found, (int)((uint32_t)__ffsll((long long)fm) - 1u));
            if (lane == sl) error_fix(J, v, ri, al, leaf, fs);
        }
    }
    wave_sync();  // the half is free again
}

// Both mates of 64 templates of a wave (one per lane): every lane walks the
// positions its reads cover (mate 0's, then mate 1's, in one loop) and queues
// (lane, mate, first record of the position) in the wave's LDS queue; the
// queue is drained 64 entries at a time, one per lane, so the per-position
// work runs on full waves instead of waiting for the lane whose reads cover
// the most positions. Sequencing errors: with DEFER the reads are counted
// with their true bases and each drained position whose observed base
// matters goes to the deferred checks (with the lane's template index rr); otherwise the
// template summaries e give the observed bases (slow templates are corrected
// afterwards). Lanes with valid == false take no part (their starts must
// still lie in the stratum).
template <bool DEFER, class TAB>
__device__ __forceinline__ void count_mates_wave(const job_dev& J, const seg_dev& S, const lds_view& v,
                                                 const TAB& tab, tmpl& t, const terr& e, bool valid,
                                                 uint32_t rr, uint32_t j, uint32_t c2, uint32_t* q, chk_ring& ck,
                                                 uint32_t lane) {
    uint32_t leaf = 0;  // the lane's source leaf, resolved once a read of it covers a record
    // allele | (x2 - x) << 8, for the drains (x2 - x < 2^13)
    const uint32_t ad = t.allele | ((t.x2 - t.x) << 8);
    // one drain: entries qb[0 .. count), one per lane
    auto drain = [&](const uint32_t* qb, uint32_t count) {
        const bool on = lane < count;
        const uint32_t qe = on ? qb[lane] : 0u;
        const uint32_t ol = qe >> 26, mate = (qe >> 25) & 1u, ri = qe & 0x1FFFFFFu;
        const uint32_t lo = (uint32_t)__shfl((int)leaf, (int)ol);
        const uint32_t ado = (uint32_t)__shfl((int)ad, (int)ol);
        const uint32_t al = ado & 0xFFu;
        const uint32_t xo = (uint32_t)__shfl((int)t.x, (int)ol) + (mate ? ado >> 8 : 0u);  // the read's start
        const uint32_t in = v.info[ri], p = v.pos[ri];
        const bool grp = !PSX_INFO_SOLO(in);
        terr eo{0u, 0u};
        const uint32_t boff = mate * J.L - xo;  // base index in the template of position p: p + boff
        if (!DEFER) eo = shfl_terr(e, ol);
        if (__ballot(on && grp)) {
            if (on && grp) count_group(J, v, ri, al, lo, eo, boff);
        }
        const bool one = on && !grp;
        const bool car = PSX_INFO_ALLELE(in) == al && carried(J, in, v.lo[ri], v.hi[ri], lo);
        const bool snv = PSX_INFO_KIND(in) == PSX_SNV;
        const uint32_t tb = snv && car ? PSX_INFO_ALT(in) : PSX_INFO_REF(in);
        const uint32_t ob = DEFER ? tb : observe(eo, p + boff, tb);
        if (one) {
            v.add_dep(ri);
            if (PSX_INFO_HEAD(in) && (snv ? ob == PSX_INFO_ALT(in) : car)) v.add_alt(ri);
        }
        if (DEFER) {
            const bool chk = J.err_n && on && (grp || (snv && PSX_INFO_HEAD(in)));  // no errors: nothing to check
            const unsigned long long m = __ballot(chk);
            if (m) {
                const uint32_t rank =
                    __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                const uint32_t ro = (uint32_t)__shfl((int)rr, (int)ol);  // the entry's template
                if (chk) {
                    uint32_t* slot = ck.e + ((ck.head + ck.count + rank) & 127u) * 3u;
                    slot[0] = ro;
                    slot[1] = (p + boff) | (al << 13) | (ri << 21);
                    slot[2] = lo;
                }
                ck.count += (uint32_t)__popcll(m);
            }
        }
    };
    uint32_t qn = 0;  // wave-uniform
    // one walk over both reads of the lane's template: mate 0's positions, then
    // mate 1's, each from its bucket start (records before the read are passed
    // over); p is the position of record i, prev that of the record before it
    const uint32_t e0 = t.x + J.L, e1 = t.x2 + J.L;
    if (v.n == 0u) return;  // at() needs n > 0
    uint32_t i = v.start(t.x);
    const uint32_t j1 = J.pe ? v.start(t.x2) : 0u;
    uint32_t p = v.pos[v.at(i)];
    const uint32_t p1 = v.pos[v.at(j1)];
    const bool go0 = valid && i < v.n && p < e0;
    const bool go1 = valid && J.pe && j1 < v.n && p1 < e1;
    if (!__ballot(go0 || go1)) return;
    if (go0 || go1) leaf = leaf_of(J, S, t);
    uint32_t mate = go0 ? 0u : 1u, end = go0 ? e0 : e1, x0 = go0 ? t.x : t.x2, prev = 0xFFFFFFFFu;
    i = go0 ? i : j1;
    p = go0 ? p : p1;
    bool go = go0 || go1;
    while (__ballot(go)) {
        const uint32_t ic = go ? i : 0u;
        const bool starts = go && p >= x0 && p != prev;
        const unsigned long long m = __ballot(starts);
        const uint32_t below =
            __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (starts) q[qn + below] = (lane << 26) | (mate << 25) | ic;
        qn += (uint32_t)__popcll(m);
        i += go ? 1u : 0u;
        const uint32_t pn = v.pos[v.at(i)];
        prev = p;
        go = go && i < v.n && pn < end;
        p = pn;
        if (!go && mate == 0u && go1) {  // mate 0 done: continue with mate 1
            mate = 1u;
            end = e1;
            x0 = t.x2;
            i = j1;
            p = p1;
            prev = 0xFFFFFFFFu;
            go = true;
        }
        if (qn >= 64u) {
            wave_sync();
            drain(q + qn - 64u, 64u);
            qn -= 64u;
            wave_sync();
            if (DEFER && ck.count >= 64u) check_errors(J, v, tab, j, c2, ck, 64u, lane);
        }
    }
    if (qn) {
        wave_sync();
        drain(q, qn);
        wave_sync();
        if (DEFER && ck.count >= 64u) check_errors(J, v, tab, j, c2, ck, 64u, lane);
    }
}

// Stratum geometry shared by both kernels.
struct stratum {
    uint32_t b, e, wend, k0, n_win;
    uint64_t nj, before;
};

__device__ __forceinline__ stratum stratum_of(const job_dev& J, const seg_dev& S, uint32_t j) {
    stratum s;
    s.b = j * PSX_STRATUM;
    s.e = min(s.b + PSX_STRATUM, S.W);
    s.before = S.T * s.b / S.W;
    s.nj = S.T * s.e / S.W - s.before;
    s.wend = min(S.len, s.e - 1 + J.reach);
    s.k0 = uniform_at(J.bucket, S.bucket_off + j);
    const uint32_t kb = min((s.wend + PSX_STRATUM - 1) / PSX_STRATUM, S.n_bucket - 1);
    s.n_win = uniform_at(J.bucket, S.bucket_off + kb) - s.k0;
    return s;
}

template <int BLOCK, int CAP>
__device__ __forceinline__ void stage_window(const job_dev& J, const stratum& s, uint32_t* s_pos, uint32_t* s_lo,
                                             uint32_t* s_hi, uint32_t* s_info, uint32_t* s_alt, uint32_t* s_dep,
                                             uint16_t* s_bucket, uint32_t tid) {
    for (uint32_t i = tid; i < s.n_win; i += BLOCK) {
        s_pos[i] = J.pos[s.k0 + i];
        s_lo[i] = J.lo[s.k0 + i];
        s_hi[i] = J.hi[s.k0 + i];
        s_info[i] = J.info[s.k0 + i];
        if (s_alt) {
            s_alt[i] = 0;
            s_dep[i] = 0;
        }
    }
    __syncthreads();
    const uint32_t nb = ((s.wend - s.b) >> NB_SHIFT) + 1;
    for (uint32_t q = tid; q < nb; q += BLOCK) {
        const uint32_t target = s.b + (q << NB_SHIFT);
        uint32_t a = 0, z = s.n_win;
        while (a < z) {
            const uint32_t mid = (a + z) >> 1;
            if (s_pos[mid] < target) a = mid + 1; else z = mid;
        }
        s_bucket[q] = (uint16_t)a;
    }
    __syncthreads();
}

// The read kernel's window: only the SNVs of [k0, k0 + n_win) that a read of
// the segment's output sample can carry (present flags), compacted in table
// order, with the bucket index over them; returns their number. Reads take
// their bases from the first carried SNV at each position, so the records
// left out (indels, records no source of the sample carries) never change a
// read. s_cnt: one word per wave.
template <int BLOCK, int CAP>
__device__ __forceinline__ uint32_t stage_snvs(const job_dev& J, const seg_dev& S, const stratum& s, uint32_t* s_pos,
                                               uint32_t* s_lo, uint32_t* s_hi, uint32_t* s_info, uint16_t* s_bucket,
                                               uint32_t* s_cnt, uint32_t tid) {
    constexpr uint32_t WAVES = BLOCK / 64;
    const uint8_t* pres = J.present + (uint64_t)S.so * J.n_mut + s.k0;
    const uint32_t wave = tid >> 6;
    uint32_t n = 0;  // block-uniform
    for (uint32_t c = 0; c < s.n_win; c += BLOCK) {
        const uint32_t i = c + tid;
        const bool in = i < s.n_win;
        const uint32_t info = in ? J.info[s.k0 + i] : 0u;
        const uint32_t pos = in ? J.pos[s.k0 + i] : 0u, lo = in ? J.lo[s.k0 + i] : 0u, hi = in ? J.hi[s.k0 + i] : 0u;
        const bool keep = in && PSX_INFO_KIND(info) == PSX_SNV && pres[i] != 0;
        const unsigned long long m = __ballot(keep);
        if ((tid & 63u) == 0) s_cnt[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = n, tot = 0;
#pragma unroll
        for (uint32_t w = 0; w < WAVES; ++w) {
            const uint32_t cw = s_cnt[w];
            off += w < wave ? cw : 0u;
            tot += cw;
        }
        if (keep) {
            const uint32_t r =
                off + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            s_pos[r] = pos;
            s_lo[r] = lo;
            s_hi[r] = hi;
            s_info[r] = info;
        }
        n += tot;
        __syncthreads();
    }
    const uint32_t nb = ((s.wend - s.b) >> NB_SHIFT) + 1;
    for (uint32_t q = tid; q < nb; q += BLOCK) {
        const uint32_t target = s.b + (q << NB_SHIFT);
        uint32_t a = 0, z = n;
        while (a < z) {
            const uint32_t mid = (a + z) >> 1;
            if (s_pos[mid] < target) a = mid + 1; else z = mid;
        }
        s_bucket[q] = (uint16_t)a;
    }
    __syncthreads();
    return n;
}

// ------------------------------------------------------------ count kernel

// A stratum whose window exceeds the LDS capacity: reads search the global
// records and count with global atomics (kept out of line so the LDS path's
// code is shared by both mates).
template <int BLOCK, bool CNA, class TAB>
__device__ __forceinline__ void count_stratum_global(const job_dev& J, const seg_dev& S, const TAB& tab,
                                                  const stratum& s, uint32_t j, uint32_t c2, uint32_t* list,
                                                  uint32_t lane, uint32_t wave, uint32_t* acc) {
    const glb_view v{J.pos + s.k0, J.lo + s.k0, J.hi + s.k0, J.info + s.k0, J.alt + S.out_off + s.k0,
                     J.depth + S.out_off + s.k0, s.n_win};
    const uint32_t nj = (uint32_t)s.nj;
    for (uint32_t base = wave * 64u; base < nj; base += BLOCK) {  // wave-uniform
        const uint32_t r = base + lane;
        bool valid = r < nj;
        tmpl t = make_template<CNA>(J, S, tab, valid ? r : base, j, s.b, s.e);
        if (CNA) {
            valid = valid && t.ok;
            *acc += valid ? 1u : 0u;
        }
        const terr e = J.err_n ? template_errors(J, tab, valid ? r : base, j, c2, valid) : terr{0u, 0u};
        bool hit = false;  // errors only matter where a read covers a record
        for (uint32_t mate = 0; mate <= J.pe; ++mate)
            hit |= count_read(J, S, v, t, e, valid, mate ? t.x2 : t.x, mate);
        if (J.err_n && hit) count_slow_errors(J, v, tab, t, e, valid, r, j, c2, list, lane);
    }
}

// CNA: some job has copy-number events (templates may be rejected and are
// counted); the plain instantiation carries none of that code. DEFER: the
// jobs' reads cover few positions per wave (sparse tables, rejected
// templates), so errors are drawn per covered position (deferred checks, one
// draw per 64 positions) instead of per template; both give the same counts.
// The deferred form is held to 6 waves per SIMD (80 VGPRs; a few spill), the
// plain form to 7 (72 VGPRs, with a 1024-entry guide so that 7 blocks of 256
// threads fit in LDS).
template <int BLOCK, int CAP, bool CNA, bool DEFER>
__global__ void __launch_bounds__(BLOCK, DEFER ? 6 : (CNA ? 1 : 7)) k_seq_count(const seq_launch L) {
    __shared__ uint32_t s_pos[CAP], s_lo[CAP], s_hi[CAP], s_info[CAP], s_alt[CAP], s_dep[CAP];
    __shared__ uint16_t s_bucket[NB_MAX];
    // the deferred form keeps a 256-entry guide (its LDS also holds the error-check rings),
    // the plain form a 1024-entry one
    __shared__ lds_tables_t<DEFER ? 8u : (CNA ? PSX_GUIDE_BITS : 10u)> s_tab;
    __shared__ uint32_t s_item, s_acc;
    // per wave: the (lane, record) queue of a mate (also a slow template's errors on
    // the global path), and the deferred error checks
    __shared__ uint32_t s_wave[BLOCK / 64][ELIST];
    __shared__ uint32_t s_chk[BLOCK / 64][DEFER ? 3 * 128 : 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    for (;;) {
        if (tid == 0) {
            s_item = (uint32_t)atomicAdd(L.queue, 1ull);
            s_acc = 0;
        }
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= L.n_items) break;
        const item_dev it = uniform_at(L.items, item);
        // the segment's and the job's fields are read where they are used (scalar
        // loads from the constant address space) instead of being held in SGPRs
        // all along, which spilled them to VGPR lanes
        const seg_dev& S = *(const seg_dev*)((const __attribute__((address_space(4))) seg_dev*)L.segs + it.seg);
        const job_dev& J = *(const job_dev*)((const __attribute__((address_space(4))) job_dev*)L.jobs + S.job);
        const uint32_t c2 = (S.chr << 16) | S.so;
        uint32_t acc = 0;  // accepted templates (copy-number events)
        load_tables(J, s_tab, tid, BLOCK);
        __syncthreads();
        for (uint32_t j = it.j0; j < it.j1; ++j) {
            const stratum s = stratum_of(J, S, j);
            if (s.n_win <= (uint32_t)CAP) {
                stage_window<BLOCK, CAP>(J, s, s_pos, s_lo, s_hi, s_info, s_alt, s_dep, s_bucket, tid);
                const lds_view v{s_pos, s_lo, s_hi, s_info, s_alt, s_dep, s_bucket, s.n_win, s.b};
                const uint32_t nj = (uint32_t)s.nj;
                chk_ring ck{s_chk[wave], 0u, 0u};
                // the counting of a wave's templates t (lanes with valid; template indices rr)
                auto count_wave = [&](tmpl& t, uint32_t rr, bool valid) {
                    if (DEFER) {
                        count_mates_wave<true>(J, S, v, s_tab, t, terr{0u, 0u}, valid, rr, j, c2, s_wave[wave], ck,
                                               lane);
                    } else {
                        const terr e = J.err_n ? template_errors(J, s_tab, rr, j, c2, valid) : terr{0u, 0u};
                        count_mates_wave<false>(J, S, v, s_tab, t, e, valid, rr, j, c2, s_wave[wave], ck, lane);
                        if (J.err_n) count_slow_errors(J, v, s_tab, t, e, valid, rr, j, c2, s_wave[wave], lane);
                    }
                };
                if (!CNA && DEFER) {
                    for (uint32_t base = wave * 64u; base < nj; base += BLOCK) {  // wave-uniform
                        const uint32_t r = base + lane;
                        const bool valid = r < nj;
                        tmpl t = make_template<CNA>(J, S, s_tab, valid ? r : base, j, s.b, s.e);
                        count_mates_wave<true>(J, S, v, s_tab, t, terr{0u, 0u}, valid, valid ? r : base, j, c2,
                                               s_wave[wave], ck, lane);
                    }
                } else if (!CNA) {
                    for (uint32_t base = wave * 64u; base < nj; base += BLOCK) {  // wave-uniform
                        const uint32_t r = base + lane;
                        const bool valid = r < nj;
                        const uint32_t rv = valid ? r : base;
                        tmpl t = make_template<CNA>(J, S, s_tab, rv, j, s.b, s.e);
                        const terr e = J.err_n ? template_errors(J, s_tab, rv, j, c2, valid) : terr{0u, 0u};
                        count_mates_wave<false>(J, S, v, s_tab, t, e, valid, rv, j, c2, s_wave[wave], ck, lane);
                        if (J.err_n) count_slow_errors(J, v, s_tab, t, e, valid, rv, j, c2, s_wave[wave], lane);
                    }
                } else {
                    // copy-number draws reject templates: the accepted ones are packed
                    // into the lanes [0, pc) of a pending set and counted once the next
                    // group would overflow it, so the walks run on nearly full waves
                    uint32_t pc = 0;  // wave-uniform
                    uint32_t px = 0, pad = 0, pleaf = 0, pr = 0;  // x; x2 - x | allele << 13; leaf; index
                    auto flush = [&]() {
                        const bool on = lane < pc;
                        tmpl t;
                        t.x = on ? px : s.b;  // idle lanes keep a start inside the stratum
                        t.x2 = on ? px + (pad & 0x1FFFu) : s.b;
                        t.allele = pad >> 13;
                        t.leaf = pleaf;
                        t.cell = 0;
                        t.ok = 1;
                        count_wave(t, pr, on);
                        pc = 0;
                    };
                    for (uint32_t base = wave * 64u; base < nj; base += BLOCK) {  // wave-uniform
                        const uint32_t r = base + lane;
                        bool valid = r < nj;
                        tmpl t = make_template<CNA>(J, S, s_tab, valid ? r : base, j, s.b, s.e);
                        valid = valid && t.ok;
                        acc += valid ? 1u : 0u;
                        const unsigned long long m = __ballot(valid);
                        const uint32_t n = (uint32_t)__popcll(m);
                        if (pc + n > 64u) flush();
                        // a permutation of the lanes: accepted ones to [pc, pc + n) in lane
                        // order, the others to the remaining lanes
                        const uint32_t below =
                            __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        const uint32_t dst = valid ? pc + below : (pc + n + (lane - below)) & 63u;
                        const int a4 = (int)(dst * 4u);
                        const uint32_t nx = (uint32_t)__builtin_amdgcn_ds_permute(a4, (int)t.x);
                        const uint32_t nad = (uint32_t)__builtin_amdgcn_ds_permute(
                            a4, (int)((t.x2 - t.x) | (t.allele << 13)));
                        const uint32_t nleaf = (uint32_t)__builtin_amdgcn_ds_permute(a4, (int)t.leaf);
                        const uint32_t nr = (uint32_t)__builtin_amdgcn_ds_permute(a4, (int)r);
                        const bool recv = lane >= pc && lane < pc + n;
                        px = recv ? nx : px;
                        pad = recv ? nad : pad;
                        pleaf = recv ? nleaf : pleaf;
                        pr = recv ? nr : pr;
                        pc += n;
                        if (pc == 64u) flush();
                    }
                    if (pc) flush();
                }
                if (DEFER && J.err_n && ck.count) check_errors(J, v, s_tab, j, c2, ck, ck.count, lane);
                __syncthreads();
                int32_t* ga = J.alt + S.out_off + s.k0;
                int32_t* gd = J.depth + S.out_off + s.k0;
                for (uint32_t i = tid; i < s.n_win; i += BLOCK) {
                    const uint32_t d = s_dep[i];
                    if (d) {
                        atomicAdd(&gd[i], (int32_t)d);
                        const uint32_t a = s_alt[i];
                        if (a) atomicAdd(&ga[i], (int32_t)a);
                    }
                }
                __syncthreads();
            } else {
                count_stratum_global<BLOCK, CNA>(J, S, s_tab, s, j, c2, s_wave[wave], lane, wave, &acc);
            }
        }
        if (CNA && S.n_cna) {
            if (acc) atomicAdd(&s_acc, acc);
            __syncthreads();
            if (tid == 0 && s_acc) atomicAdd(J.acc, (unsigned long long)s_acc * (J.pe + 1u));
        }
        __syncthreads();
    }
}

// Presence (the sample holds the mutation) and zeroed counters; also resets
// the work counter of the count kernel that follows on the stream.
// Whether record m can be carried by a read of output sample P: a leaf of the
// sample lies in its leaf range, or the normal genome carries it and the
// sample's reads can come from the normal genome (the normal sample, or purity
// below 1). Oracle: the presence flags of orc_simulate_seq.
__device__ __forceinline__ bool record_present(const job_dev& J, const pres_dev& P, uint64_t m) {
    const uint32_t info = J.info[m];
    const uint32_t nat = PSX_INFO_NATURE(info);
    const bool normal_has = nat == PSX_GERMINAL || (nat == PSX_PRENEOPLASTIC && J.preneo);
    bool pres;
    if (P.n_cells == 0) {
        pres = normal_has;
    } else {
        pres = J.purity_lt1 && normal_has;
        if (!pres) {
            // sample leaves of rank below x, from the sample's running counts
            const uint4* lc = J.leaf_count + (uint64_t)P.sample * J.leaf_blocks;
            auto below = [&](uint32_t x) {
                const uint4 e = lc[x >> 6];
                const uint64_t bits = (uint64_t)e.y | ((uint64_t)e.z << 32);
                return e.x + (uint32_t)__popcll(bits & ((1ull << (x & 63u)) - 1ull));
            };
            pres = below(J.hi[m]) > below(J.lo[m]);
        }
    }
    return pres;
}

// The (output sample, mutation) pair of thread q: ps lists each pair range's
// first index.
__device__ __forceinline__ const pres_dev& pres_of(const pres_dev* __restrict__ ps, uint32_t n_ps, uint64_t q) {
    uint32_t a = 0, z = n_ps;
    while (a < z) {
        const uint32_t mid = (a + z) >> 1;
        if (ps[mid].begin <= q) a = mid + 1; else z = mid;
    }
    return ps[a - 1];
}

__global__ void __launch_bounds__(256) k_seq_init(const job_dev* __restrict__ jobs, const pres_dev* __restrict__ ps,
                                                  uint32_t n_ps, uint64_t total,
                                                  unsigned long long* __restrict__ queue, uint32_t n_jobs) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) {
        queue[0] = 0ull;  // the sparse jobs' item counter
        queue[1] = 0ull;  // the others'
        for (uint32_t k = 0; k < n_jobs; ++k)
            if (jobs[k].acc) *jobs[k].acc = 0ull;
    }
    if (q >= total) return;
    const pres_dev P = pres_of(ps, n_ps, q);
    const job_dev& J = jobs[P.job];
    const uint64_t m = q - P.begin;
    const uint64_t o = (uint64_t)P.so * J.n_mut + m;
    J.alt[o] = 0;
    J.depth[o] = 0;
    J.present[o] = record_present(J, P, m) ? 1 : 0;
}

// Presence flags only (read materialisation: the read kernel stages the SNVs
// a read of the sample can carry).
__global__ void __launch_bounds__(256) k_reads_present(const job_dev* __restrict__ jobs,
                                                       const pres_dev* __restrict__ ps, uint32_t n_ps, uint64_t total) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= total) return;
    const pres_dev P = pres_of(ps, n_ps, q);
    const job_dev& J = jobs[P.job];
    const uint64_t m = q - P.begin;
    J.present[(uint64_t)P.so * J.n_mut + m] = record_present(J, P, m) ? 1 : 0;
}

// --------------------------------------------------- read materialisation
//
// Work layout. A wave takes 64 templates of the stratum at a time (one per
// lane: draws, insert, source leaf), then writes their reads in passes of 64
// reads, one read per lane, in output order: with paired ends lane l of pass
// p holds mate l % 2 of template p * 32 + l / 2, whose fields it fetches from
// the lane that drew it. A read is built in the lane's row of a wave-private
// LDS stage and each pass is copied out by the whole wave with 16-byte
// coalesced stores, so no block barrier separates generation from the copy.

// One read into row[0 .. 2 + ceil(L/16)): word 0 = start, word 1 = leaf |
// allele << 26 | mate << 30 | (errors > 0) << 31, then the 2-bit bases (base
// i of the read in bits 2 (i % 16) of word 2 + i / 16) taken from the staged
// reference (s_ref[0] is reference word ref_b / 16), with the first carried
// SNV at each position and the template summary's errors of this mate
// applied; bases past the read end are 0. A slow template's errors are
// applied to the row afterwards (apply_error). Oracle: read_row.
template <int NW, class V>
__device__ __forceinline__ void write_read(const job_dev& J, const V& v, uint32_t x, uint32_t leaf, uint32_t allele,
                                           uint32_t mate, const terr& e, const uint32_t* s_ref, uint32_t ref_b,
                                           uint32_t* row) {
    const uint32_t tail = J.L & 15u;
    const uint32_t xr = x - ref_b;
    const uint32_t w = xr >> 4;
    const uint32_t sh = (xr & 15u) * 2u;
    // s_ref holds one word past the staged range, so w + k + 1 is always readable
    if constexpr (NW > 0) {
        // 16-byte row: bases built in registers, stored as three 128-bit writes
        uint32_t wd[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) wd[k] = __builtin_amdgcn_alignbit(s_ref[w + k + 1], s_ref[w + k], sh);
        if (tail) wd[NW - 1] &= (1u << (2u * tail)) - 1u;
        static_assert((NW + 2) % 4 == 0, "rows of whole 16-byte vectors");
        uint4* r4 = reinterpret_cast<uint4*>(row);
        r4[0] = uint4{x, 0u, wd[0], wd[1]};
#pragma unroll
        for (int q = 1; q < (NW + 2) / 4; ++q) r4[q] = uint4{wd[4 * q - 2], wd[4 * q - 1], wd[4 * q], wd[4 * q + 1]};
    } else {
        const uint32_t nw = (J.L + 15u) >> 4;
        uint32_t lo = s_ref[w];
        for (uint32_t k = 0; k < nw; ++k) {
            const uint32_t hi = s_ref[w + k + 1];
            uint32_t word = __builtin_amdgcn_alignbit(hi, lo, sh);  // bits [sh, sh + 32) of hi:lo
            if (k + 1 == nw && tail) word &= (1u << (2u * tail)) - 1u;
            row[2 + k] = word;
            lo = hi;
        }
    }
    // SNVs at [x, x + L) in table order; the first carried SNV at a position
    // sets its base. One loop from the lane's bucket start (records before x
    // are passed over), uniform trip count, branch-free read-modify-writes of
    // the lane's own row (an unchanged word when nothing applies).
    const uint32_t end = x + J.L;
    uint32_t i = v.start(x);
    uint32_t p = v.n ? v.pos[v.at(i)] : end;  // at() needs n > 0
    bool go = i < v.n && p < end;
    uint32_t last = 0xFFFFFFFFu;
    while (__ballot(go)) {
        const uint32_t ic = go ? i : 0u;
        const uint32_t in = v.info[ic], lo = v.lo[ic], hi = v.hi[ic];
        i += go ? 1u : 0u;
        const uint32_t pn = v.pos[v.at(i)];
        const bool apply = go && p >= x && p != last && PSX_INFO_KIND(in) == PSX_SNV &&
                           PSX_INFO_ALLELE(in) == allele && carried(J, in, lo, hi, leaf);
        if (apply) {  // rare: only applying lanes touch their row (no bank conflicts from the others)
            const uint32_t o = p - x, s2 = (o & 15u) * 2u;
            row[2 + (o >> 4)] = (row[2 + (o >> 4)] & ~(3u << s2)) | (PSX_INFO_ALT(in) << s2);
        }
        last = apply ? p : last;
        go = go && i < v.n && pn < end;
        p = pn;
    }
    // the summary's errors on this mate (wave-uniform trip count, predicated)
    const uint32_t ne = e.n();
    uint32_t flag = 0;
    for (uint32_t q = 0; __ballot(q < ne); ++q) {
        const uint32_t eq = q == 0u ? (e.w01 & 0xFFFFu) : (q == 1u ? e.w01 >> 16 : e.w2n & 0xFFFFu);
        const uint32_t b = eq & 0x1FFFu;
        const uint32_t m = b >= J.L ? 1u : 0u;
        if (q < ne && m == mate) {
            const uint32_t o = b - m * J.L, s2 = (o & 15u) * 2u;
            const uint32_t wv = row[2 + (o >> 4)];
            const uint32_t tb = (wv >> s2) & 3u, sub = eq >> 13;
            const uint32_t nb = sub < tb ? sub : sub + 1u;
            row[2 + (o >> 4)] = (wv & ~(3u << s2)) | (nb << s2);
            flag = 1u;
        }
    }
    if constexpr (NW == 0) row[0] = x;
    row[1] = leaf | (allele << 26) | (mate << 30) | (flag << 31);
}

// A sequencing error at offset o of a written read: the base there is
// replaced by substitution index sub among the three other bases, and the
// read is flagged. ATOMIC: the row is shared with other lanes that may patch
// other bases of the same word at the same time (no two errors share a base,
// so the base read here is stable).
template <bool ATOMIC>
__device__ __forceinline__ void apply_error(uint32_t* row, uint32_t o, uint32_t sub) {
    uint32_t* wp = row + 2 + (o >> 4);
    const uint32_t s2 = (o & 15u) * 2u;
    const uint32_t tb = (*wp >> s2) & 3u;
    const uint32_t nb = sub < tb ? sub : sub + 1u;
    if (ATOMIC) {
        atomicXor(wp, (tb ^ nb) << s2);
        atomicOr(row + 1, 0x80000000u);
    } else {
        *wp ^= (tb ^ nb) << s2;
        row[1] |= 0x80000000u;
    }
}

// STAGE: reads are built in LDS and copied out per pass (read_words small
// enough for 64 rows per wave); otherwise each lane builds its read in place
// in the output. NW > 0: reads of NW base words (rows of whole 16-byte
// vectors, built in registers); NW = 0: any read size. CNA: segments with
// copy-number events keep only the templates whose copy slot is a copy of
// their source (their reads from J.cn_row on); the others as without.
template <int BLOCK, int CAP, bool STAGE, int NW, bool CNA = false>
__global__ void __launch_bounds__(BLOCK) k_seq_reads(const seq_launch L) {
    constexpr uint32_t WAVES = BLOCK / 64;
    __shared__ uint32_t s_pos[CAP], s_lo[CAP], s_hi[CAP], s_info[CAP];
    __shared__ uint16_t s_bucket[NB_MAX];
    __shared__ uint32_t s_ref[REF_MAX];
    __shared__ lds_tables s_tab;
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_cnt[WAVES];
    __shared__ uint32_t s_elist[WAVES][ELIST];  // per wave: errors of a slow template
    __shared__ uint32_t s_kept[WAVES];          // CNA: kept templates per wave in a round
    extern __shared__ uint4 s_stage4[];  // STAGE: per wave 64 rows of read_words words
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    uint32_t* list = s_elist[wave];
    const uint32_t RW = L.read_words;
    const bool vec4 = L.out_vec4 != 0;  // RW % 4 == 0 and a 16-byte aligned output
    uint32_t* stage = reinterpret_cast<uint32_t*>(s_stage4) + wave * 64u * RW;
    const uint4* stage4 = reinterpret_cast<const uint4*>(stage);
    for (;;) {
        if (tid == 0) s_item = (uint32_t)atomicAdd(L.queue, 1ull);
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= L.n_items) break;
        const item_dev it = uniform_at(L.items, item);
        const seg_dev S = uniform_at(L.segs, it.seg);
        const job_dev J = uniform_at(L.jobs, S.job);
        const uint32_t c2 = (S.chr << 16) | S.so;
        const uint32_t rpt = J.pe ? 2u : 1u;
        const uint32_t tpp = 64u / rpt;  // templates per pass
        const uint32_t chr_words = (S.len + 15u) >> 4;
        load_tables(J, s_tab, tid, BLOCK);
        __syncthreads();
        for (uint32_t j = it.j0; j < it.j1; ++j) {
            const stratum s = stratum_of(J, S, j);
            const bool fits = s.n_win <= (uint32_t)CAP;
            // reference words covering [b, wend) plus one
            const uint32_t ref_w0 = s.b >> 4;
            const uint32_t n_ref = min(chr_words, ((s.wend + 15u) >> 4) + 1u) - ref_w0;
            for (uint32_t i = tid; i <= n_ref; i += BLOCK) s_ref[i] = i < n_ref ? J.packed[S.chr_woff + ref_w0 + i] : 0u;
            uint32_t n_snv = 0;
            if (fits) n_snv = stage_snvs<BLOCK, CAP>(J, S, s, s_pos, s_lo, s_hi, s_info, s_bucket, s_cnt, tid);
            else __syncthreads();
            const lds_view lv{s_pos, s_lo, s_hi, s_info, nullptr, nullptr, s_bucket, n_snv, s.b};
            const glb_view gv{J.pos + s.k0, J.lo + s.k0, J.hi + s.k0, J.info + s.k0, nullptr, nullptr, s.n_win};
            const uint32_t nj = (uint32_t)s.nj;
            // Rounds of WAVES x 64 templates, 64 per wave. Without copy-number events wave w takes templates
            // [base + 64 w, base + 64 w + 64). With them only the templates whose copy slot is a copy of their
            // source have reads, in template order: each wave counts its kept templates, one block barrier
            // orders the waves, and each wave moves its kept templates to its first lanes (forward lane
            // permutation). Either way lanes [0, n_t) then hold the templates (start tx, mate start tx2,
            // leaf | allele << 26, template index tr) whose reads start at read index first0.
            const bool cn = CNA && S.n_cna != 0u;
            const uint64_t row0 = cn ? uniform_at(J.cn_row, S.st_off + j) : 0ull;
            // Without copy-number events the chunks of 64 templates are shifted by o so that every pass after
            // the stratum's first chunk starts on a 1024-byte boundary of the output (else 512, 256, 128): the
            // first chunk holds the templates before the first aligned row. A 16-byte offset of such 3 KiB
            // passes costs about 10 % of the write rate (partly written lines at every pass boundary).
            uint32_t o = 0;
            if (!cn && STAGE && vec4) {
                const uint64_t a0 = (uint64_t)(uintptr_t)(L.reads_out + (S.read_base + s.before * rpt) * RW);
                const uint64_t step = (uint64_t)rpt * RW * 4u;  // bytes per template
                for (uint32_t A = 1024u; A >= 128u; A >>= 1) {
                    const unsigned long long m = __ballot(((a0 + lane * step) & (A - 1u)) == 0u);
                    if (m) {
                        o = (64u - ((uint32_t)__ffsll((long long)m) - 1u)) & 63u;
                        break;
                    }
                }
            }
            uint32_t kept = 0;  // copy-number events: kept templates of the stratum's earlier rounds
            for (uint32_t base = 0; base < nj + o; base += WAVES * 64u) {
                // the wave's chunk: templates [c0, cend) (virtual templates [vb, vb + 64) less o)
                const uint32_t vb = base + wave * 64u;
                const uint32_t c0 = vb == 0u ? 0u : vb - o;
                const uint32_t cend = min(vb + 64u - o, nj);
                const uint32_t rt = c0 + lane;
                uint32_t tx = 0, tx2 = 0, tla = 0, tr = rt, n_t;
                uint64_t first0;
                if (cn) {
                    tmpl t{};
                    bool ok = false;
                    if (rt < nj) {
                        t = make_template<true>(J, S, s_tab, rt, j, s.b, s.e);
                        ok = t.ok != 0u;
                    }
                    const unsigned long long bal = __ballot(ok);
                    if (lane == 0) s_kept[wave] = (uint32_t)__popcll(bal);
                    __syncthreads();
                    uint32_t before = kept, total = 0;
                    for (uint32_t w2 = 0; w2 < WAVES; ++w2) {
                        const uint32_t c = s_kept[w2];
                        before += w2 < wave ? c : 0u;
                        total += c;
                    }
                    __syncthreads();  // every wave has read s_kept before the next round writes it
                    kept += total;
                    n_t = (uint32_t)__popcll(bal);
                    const uint32_t mine = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                    const int to = (int)(ok ? mine : n_t + (lane - mine)) << 2;
                    tla = leaf_of(J, S, t) | (t.allele << 26);
                    tx = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)t.x);
                    tx2 = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)t.x2);
                    tla = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)tla);
                    tr = (uint32_t)__builtin_amdgcn_ds_permute(to, (int)rt);
                    first0 = row0 + (uint64_t)before * rpt;
                } else {
                    n_t = cend > c0 ? cend - c0 : 0u;
                    if (lane < n_t) {
                        tmpl t = make_template(J, S, s_tab, rt, j, s.b, s.e);
                        tx = t.x;
                        tx2 = t.x2;
                        tla = leaf_of(J, S, t) | (t.allele << 26);
                    }
                    first0 = S.read_base + (s.before + c0) * rpt;
                }
                if (n_t == 0) continue;
                const bool valid = lane < n_t;
                const uint32_t tr0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)tr);  // lane 0 holds a template
                const terr te = J.err_n ? template_errors(J, s_tab, valid ? tr : tr0, j, c2, valid) : terr{0u, 0u};
                for (uint32_t p = 0; p < rpt; ++p) {
                    const uint32_t src = p * tpp + lane / rpt;
                    const uint32_t mate = lane & (rpt - 1u);
                    const uint32_t x1 = (uint32_t)__shfl((int)tx, (int)src);
                    const uint32_t x2 = (uint32_t)__shfl((int)tx2, (int)src);
                    const uint32_t la = (uint32_t)__shfl((int)tla, (int)src);
                    const terr e = shfl_terr(te, src);
                    const uint32_t t0 = p * tpp;  // first template of the pass
                    const uint32_t nt = t0 < n_t ? min(tpp, n_t - t0) : 0u;
                    const uint64_t first = first0 + (uint64_t)t0 * rpt;
                    uint32_t* row = STAGE ? stage + lane * RW : L.reads_out + (first + lane) * RW;
                    if (src < n_t) {
                        const uint32_t x = mate ? x2 : x1;
                        if (fits)
                            write_read<NW>(J, lv, x, la & 0x3FFFFFFu, la >> 26, mate, e, s_ref, ref_w0 << 4, row);
                        else
                            write_read<NW>(J, gv, x, la & 0x3FFFFFFu, la >> 26, mate, e, s_ref, ref_w0 << 4, row);
                    }
                    // slow templates of the pass (template lanes [t0, t0 + nt)): the wave generates each
                    // one's errors and patches its reads
                    unsigned long long sm = __ballot(te.slow() && lane >= t0 && lane < t0 + nt);
                    while (sm) {  // wave-uniform, rare
                        const uint32_t sl = (uint32_t)__ffsll((long long)sm) - 1u;
                        sm &= sm - 1ull;
                        // the slow template's index (contiguous without copy-number events)
                        const uint32_t rs = cn ? (uint32_t)__builtin_amdgcn_readlane((int)tr, (int)sl) : tr0 + sl;
                        wave_sync();
                        const uint32_t n_err = wave_template_errors(J.seed, J.err_n, J.err_k0, J.err_N, s_tab.err,
                                                                    rs, j, c2, list, lane);
                        wave_sync();  // list entries and (STAGE) every row of the pass written
                        for (uint32_t k0 = 0; k0 < n_err; k0 += 64u) {
                            const uint32_t k = k0 + lane;
                            const uint32_t w = k < n_err ? list[k] : 0u;
                            const uint32_t b = w & 0xFFFFFu, sub = w >> 20;
                            const uint32_t m = b >= J.L ? 1u : 0u;
                            const uint32_t o = b - m * J.L;
                            const uint32_t rl = (sl - t0) * rpt + m;  // the read's lane in the pass
                            const bool on = k < n_err;
                            if (STAGE) {
                                if (on) apply_error<true>(stage + rl * RW, o, sub);
                            } else {
                                // rows live in global memory: each lane patches its own read
                                for (uint32_t q = 0; q < 64u && k0 + q < n_err; ++q) {
                                    const uint32_t rq = (uint32_t)__shfl((int)rl, (int)q);
                                    const uint32_t oq = (uint32_t)__shfl((int)o, (int)q);
                                    const uint32_t sq = (uint32_t)__shfl((int)sub, (int)q);
                                    if (rq == lane) apply_error<false>(row, oq, sq);
                                }
                            }
                        }
                    }
                    if (STAGE) {
                        wave_sync();
                        const uint32_t words = nt * rpt * RW;
                        uint32_t* dst = L.reads_out + first * RW;
                        if (vec4) {
                            // plain stores: non-temporal ones cap this write pattern near 4.5 TB/s
                            uint4* d4 = reinterpret_cast<uint4*>(dst);
                            for (uint32_t q = lane; q < (words >> 2); q += 64u) d4[q] = stage4[q];
                        } else {
                            for (uint32_t q = lane; q < words; q += 64u) dst[q] = stage[q];
                        }
                        wave_sync();
                    }
                }
            }
            __syncthreads();
        }
        __syncthreads();
    }
}

// Kept templates of every stratum of the segments with copy-number events
// (the read kernel's draws, so its row order follows): one wave per stratum,
// 64 templates at a time.
struct accept_args {
    const job_dev* jobs;
    const seg_dev* segs;
    const uint32_t* cn_seg;    // segment index of each segment with events
    const uint32_t* cn_first;  // its first stratum (index into kept), n_seg + 1 entries
    uint32_t n_seg;
    uint64_t* kept;            // per stratum
};

__global__ void __launch_bounds__(256) k_cn_accept(const accept_args a) {
    const uint32_t g = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    const uint32_t lane = threadIdx.x & 63u;
    if (g >= uniform_at(a.cn_first, a.n_seg)) return;  // the whole wave
    uint32_t lo = 0, hi = a.n_seg;  // the segment: the last k with cn_first[k] <= g
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uniform_at(a.cn_first, mid) <= g) lo = mid; else hi = mid;
    }
    const seg_dev S = uniform_at(a.segs, uniform_at(a.cn_seg, lo));
    const job_dev J = uniform_at(a.jobs, S.job);
    const uint32_t j = g - uniform_at(a.cn_first, lo);
    const uint32_t b = j * PSX_STRATUM, e = min(b + PSX_STRATUM, S.W);
    const uint64_t nj = S.T * e / S.W - S.T * b / S.W;
    uint64_t cnt = 0;
    for (uint64_t r0 = 0; r0 < nj; r0 += 64u) {
        const uint64_t r = r0 + lane;
        bool ok = false;
        if (r < nj) {
            psx_u4 w;
            ok = template_source<true>(J, S, (uint32_t)r, j, b, e, w).ok != 0u;
        }
        cnt += (uint64_t)__popcll(__ballot(ok));
    }
    if (lane == 0) a.kept[g] = cnt;
}

}  // namespace psxk
