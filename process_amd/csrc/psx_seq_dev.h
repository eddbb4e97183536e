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
// can reach (pos, leaf range, packed info) with a 64-base bucket index, plus
// the job's inversion tables; reads are generated from counters and looked up
// in LDS, counts go to LDS counters and the block flushes them with one global
// atomic per touched record. All counting is integer.
#pragma once
#include "psx_cna.h"
#include "psx_internal.h"

namespace psxk {

constexpr uint32_t NB_SHIFT = 6;      // 64-base LDS buckets
constexpr uint32_t MAX_READ = 4096;   // read size and insert reach limit
constexpr uint32_t NB_MAX = (PSX_STRATUM + MAX_READ) / 64 + 2;
constexpr uint32_t MAX_INS = 640;     // compact insert-size thresholds in LDS
constexpr uint32_t MAX_ERRT = PSX_MAX_READ_ERRORS + 1;
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
    int32_t* alt;
    int32_t* depth;
    uint8_t* present;
    const uint32_t* ins_T;       // compact insert-size thresholds
    const uint16_t* ins_guide;
    const uint32_t* err_T;       // compact error-count thresholds (first value 0)
    const uint32_t* packed;      // reference (read materialisation)
    uint64_t n_mut;
    uint64_t thr_normal;         // floor((1 - purity) 2^32)
    uint32_t ins_k0, ins_n, err_n;
    uint32_t L, pe, ploidy, n_leaves, preneo, reach;
    uint32_t n_out;
    uint32_t purity_lt1;
    const psx_cna_dev* cna;      // copy-number events of every chromosome, per segment ranges
    unsigned long long* acc;     // reads of accepted templates (jobs with copy-number events)
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
    uint32_t kmax, pad;          // copies drawn per template: the largest copy number, >= ploidy
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
    uint32_t out_vec4;       // read_words % 4 == 0 and reads_out 16-byte aligned
};

struct lds_tables {
    uint32_t ins[MAX_INS];
    uint16_t guide[PSX_GUIDE];
    uint32_t err[MAX_ERRT];
};

__device__ __forceinline__ void load_tables(const job_dev& J, lds_tables& t, uint32_t tid, uint32_t block) {
    if (J.pe) {
        for (uint32_t i = tid; i < J.ins_n; i += block) t.ins[i] = J.ins_T[i];
        for (uint32_t i = tid; i < PSX_GUIDE; i += block) t.guide[i] = J.ins_guide[i];
    }
    for (uint32_t i = tid; i < J.err_n; i += block) t.err[i] = J.err_T[i];
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
    // Same result; the wave steps together (uniform trip count, no exec
    // masking), every load in bounds.
    __device__ __forceinline__ uint32_t first_u(uint32_t x) const {
        uint32_t i = bucket[(x - b) >> NB_SHIFT];
        bool go = i < n && pos[min(i, n - 1u)] < x;
        while (__ballot(go)) {
            i += go ? 1u : 0u;
            go = go && i < n && pos[min(i, n - 1u)] < x;
        }
        return i;
    }
    __device__ __forceinline__ uint32_t at(uint32_t i) const { return min(i, n - 1u); }  // n > 0
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
    __device__ __forceinline__ uint32_t first_u(uint32_t x) const { return first(x); }
    __device__ __forceinline__ uint32_t at(uint32_t i) const { return min(i, n - 1u); }  // n > 0
    __device__ __forceinline__ void add_dep(uint32_t q) const { atomicAdd(&dep[q], 1); }
    __device__ __forceinline__ void add_alt(uint32_t q) const { atomicAdd(&alt[q], 1); }
};

// Template: start of mate 1 and mate 2, source cell (resolved to a leaf on
// first use) and copy; ok == 0 when a copy-number draw rejects it.
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

// Template r of stratum j (oracle: the template loop of orc_simulate_seq).
template <bool CNA = false>
__device__ __forceinline__ tmpl make_template(const job_dev& J, const seg_dev& S, const lds_tables& tab, uint32_t r,
                                              uint32_t j, uint32_t b, uint32_t e) {
    const uint32_t c2 = (S.chr << 16) | S.so;
    const psx_u4 w = psx_draw_u(J.seed, r, j, c2, PSX_P_TEMPLATE << 24);
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
            t.leaf = UNRESOLVED;
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
    t.x2 = t.x;
    if (J.pe) {
        uint32_t I = J.ins_k0 + psx_guided_pick(tab.ins, tab.guide, J.ins_n, w.w);
        if (I < J.L) I = J.L;
        uint64_t x2 = (uint64_t)t.x + I - J.L;
        if (x2 > S.len - J.L) x2 = S.len - J.L;
        t.x2 = (uint32_t)x2;
    }
    return t;
}

// Error set of one read (oracle: error_set): K ~ Binomial(L, e) from word 0,
// then K distinct offsets, one word each; a repeated offset takes the next
// word. Slots hold offset | substitution << 12.
__device__ uint32_t error_set(const job_dev& J, const lds_tables& tab, uint32_t r, uint32_t j, uint32_t c2,
                              uint32_t mate, uint16_t* slots, uint32_t stride) {
    const uint32_t base = (PSX_P_ERROR << 24) | (mate << 16);
    psx_u4 w = psx_draw_u(J.seed, r, j, c2, base);
    uint32_t K = 0;
    while (K + 1 < J.err_n && w.x >= tab.err[K]) ++K;
    uint32_t cnt = 0;
    for (uint32_t t = 1; cnt < K && t < 1 + 4 * PSX_MAX_READ_ERRORS; ++t) {
        if ((t & 3u) == 0) w = psx_draw_u(J.seed, r, j, c2, base | (t >> 2));
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

// error_set with wave-uniform loops: every lane steps through the same
// draws (the word counter t is uniform), effects are predicated, so the
// draw sequence and the result of each lane are those of error_set.
// Lanes with want == false return 0 without stepping the loops.
__device__ __forceinline__ uint32_t error_set_u(const job_dev& J, const lds_tables& tab, uint32_t r, uint32_t j,
                                                uint32_t c2, uint32_t mate, uint16_t* slots, uint32_t stride,
                                                bool want = true) {
    const uint32_t base = (PSX_P_ERROR << 24) | (mate << 16);
    psx_u4 w = psx_draw_u(J.seed, r, j, c2, base);
    uint32_t K = 0;
    bool more = want && 1u < J.err_n && w.x >= tab.err[0];
    while (__ballot(more)) {
        K += more ? 1u : 0u;
        more = more && K + 1u < J.err_n && w.x >= tab.err[min(K, J.err_n - 1u)];
    }
    uint32_t cnt = 0;
    bool need = K > 0;
    for (uint32_t t = 1; t < 1u + 4u * PSX_MAX_READ_ERRORS && __ballot(need); ++t) {
        if ((t & 3u) == 0) w = psx_draw_u(J.seed, r, j, c2, base | (t >> 2));
        const uint32_t word = psx_word(w, (int)(t & 3u));
        const uint32_t o = psx_mulhi32(word, J.L);
        bool dup = false;
        const uint32_t qn = min(t - 1u, PSX_MAX_READ_ERRORS);
        for (uint32_t q = 0; q < qn; ++q) dup |= q < cnt && (slots[q * stride] & 0xFFFu) == o;
        const bool take = need && !dup;
        if (take) slots[cnt * stride] = (uint16_t)(o | ((((word & 0xFFFFu) * 3u) >> 16) << 12));
        cnt += take ? 1u : 0u;
        need = cnt < K;
    }
    return cnt;
}

// Records of one position shared by several records, from index i (the
// oracle's seq_read step): the first carried SNV sets the true base, the
// read's error at the position substitutes it, every record gains depth and
// each SID head an alternative read when its alternative is observed (SNV)
// or carried (indel). Returns the index past the position.
template <class V>
__device__ __forceinline__ uint32_t count_group(const job_dev& J, const V& v, uint32_t i, uint32_t x, uint32_t allele,
                                             uint32_t leaf, uint32_t ne, const uint16_t* slots, uint32_t stride) {
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
    uint32_t ob = tb;
    for (uint32_t q = 0; q < ne; ++q) {
        const uint32_t e = slots[q * stride];
        if ((e & 0xFFFu) == p - x) {
            const uint32_t sub = e >> 12;
            ob = sub < tb ? sub : sub + 1;
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
            for (uint32_t q2 = q; q2 < g1 && (q2 == q || !PSX_INFO_HEAD(v.info[q2])); ++q2) {
                const uint32_t in2 = v.info[q2];
                car |= PSX_INFO_ALLELE(in2) == allele && carried(J, in2, v.lo[q2], v.hi[q2], leaf);
            }
            if (car) v.add_alt(q);
        }
    }
    return g1;
}

// One read [x, x + L): depth for every record at a covered position, an
// alternative read for each SID head whose alternative is observed (SNV) or
// carried (indel). Oracle: seq_read. A record alone at its position (every
// record under the infinite-sites model) takes one branch-free step of a
// wave-uniform loop; a position shared by several records goes through
// count_group.
template <class V>
__device__ __forceinline__ void count_read(const job_dev& J, const seg_dev& S, const V& v, const lds_tables& tab,
                                           tmpl& t, uint32_t x, uint32_t r, uint32_t j, uint32_t c2, uint32_t mate,
                                           uint16_t* slots, uint32_t stride) {
    const uint32_t end = x + J.L;
    uint32_t i = v.first_u(x);
    bool go = i < v.n && v.pos[v.at(i)] < end;
    if (!__ballot(go)) return;
    const uint32_t leaf = leaf_of(J, S, t);
    const uint32_t ne = J.err_n ? error_set_u(J, tab, r, j, c2, mate, slots, stride, go) : 0u;
    while (__ballot(go)) {
        const uint32_t ic = go ? i : 0u;
        const uint32_t p = v.pos[ic], in = v.info[ic];
        if (__ballot(go && !PSX_INFO_SOLO(in))) {
            if (go && !PSX_INFO_SOLO(in)) i = count_group(J, v, ic, x, t.allele, leaf, ne, slots, stride);
        }
        const bool one = go && PSX_INFO_SOLO(in);
        const bool car = PSX_INFO_ALLELE(in) == t.allele && carried(J, in, v.lo[ic], v.hi[ic], leaf);
        const bool snv = PSX_INFO_KIND(in) == PSX_SNV;
        const uint32_t tb = snv && car ? PSX_INFO_ALT(in) : PSX_INFO_REF(in);
        uint32_t ob = tb;
        for (uint32_t q = 0; __ballot(one && q < ne); ++q) {
            const bool on = one && q < ne;
            const uint32_t e = on ? slots[q * stride] : 0u;
            const uint32_t sub = e >> 12;
            ob = on && (e & 0xFFFu) == p - x ? (sub < tb ? sub : sub + 1) : ob;
        }
        if (one) {
            v.add_dep(ic);
            if (PSX_INFO_HEAD(in) && (snv ? ob == PSX_INFO_ALT(in) : car)) v.add_alt(ic);
        }
        i += one ? 1u : 0u;
        go = go && i < v.n && v.pos[v.at(i)] < end;
    }
}

// One mate of 64 templates of a wave (one per lane): every lane finds the
// positions its read covers and queues (lane, first record of the position)
// in the wave's LDS queue; the queue is drained 64 entries at a time, one
// per lane, so the per-position work runs on full waves instead of waiting
// for the lane whose read covers the most positions. Same counts as
// count_read. Lanes with valid == false take no part (x must still lie in
// the stratum).
template <int BLOCK>
__device__ __forceinline__ void count_mate_wave(const job_dev& J, const seg_dev& S, const lds_view& v,
                                                const lds_tables& tab, tmpl& t, bool valid, uint32_t x, uint32_t r,
                                                uint32_t j, uint32_t c2, uint32_t mate, uint16_t* s_err,
                                                uint32_t* q, uint32_t lane, uint32_t wave) {
    const uint32_t end = x + J.L;
    uint32_t i = v.first_u(x);
    bool go = valid && i < v.n && v.pos[v.at(i)] < end;
    if (!__ballot(go)) return;
    uint32_t leaf = 0;
    if (go) leaf = leaf_of(J, S, t);
    uint16_t* wslots = s_err + wave * 64u;  // slot q of lane l: wslots[l + q * BLOCK]
    const uint32_t ne = J.err_n ? error_set_u(J, tab, r, j, c2, mate, wslots + lane, BLOCK, go) : 0u;
    const uint32_t an = t.allele | (ne << 8);
    // one drain: entries qb[0 .. count), one per lane
    auto drain = [&](const uint32_t* qb, uint32_t count) {
        const bool on = lane < count;
        const uint32_t e = on ? qb[lane] : 0u;
        const uint32_t ol = e >> 26, ri = e & 0x3FFFFFFu;
        const uint32_t xo = (uint32_t)__shfl((int)x, (int)ol);
        const uint32_t lo = (uint32_t)__shfl((int)leaf, (int)ol);
        const uint32_t ao = (uint32_t)__shfl((int)an, (int)ol);
        const uint32_t al = ao & 0xFFu, neo = ao >> 8;
        const uint16_t* sl = wslots + ol;
        const uint32_t in = v.info[ri], p = v.pos[ri];
        if (__ballot(on && !PSX_INFO_SOLO(in))) {
            if (on && !PSX_INFO_SOLO(in)) count_group(J, v, ri, xo, al, lo, neo, sl, BLOCK);
        }
        const bool one = on && PSX_INFO_SOLO(in);
        const bool car = PSX_INFO_ALLELE(in) == al && carried(J, in, v.lo[ri], v.hi[ri], lo);
        const bool snv = PSX_INFO_KIND(in) == PSX_SNV;
        const uint32_t tb = snv && car ? PSX_INFO_ALT(in) : PSX_INFO_REF(in);
        uint32_t ob = tb;
        for (uint32_t k = 0; __ballot(one && k < neo); ++k) {
            const bool hit = one && k < neo;
            const uint32_t ew = hit ? sl[k * BLOCK] : 0u;
            const uint32_t sub = ew >> 12;
            ob = hit && (ew & 0xFFFu) == p - xo ? (sub < tb ? sub : sub + 1) : ob;
        }
        if (one) {
            v.add_dep(ri);
            if (PSX_INFO_HEAD(in) && (snv ? ob == PSX_INFO_ALT(in) : car)) v.add_alt(ri);
        }
    };
    const uint32_t i0 = i;
    uint32_t qn = 0;  // wave-uniform
    while (__ballot(go)) {
        const uint32_t ic = go ? i : 0u;
        const bool starts = go && (i == i0 || v.pos[ic] != v.pos[ic ? ic - 1u : 0u]);
        const unsigned long long m = __ballot(starts);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (starts) q[qn + below] = (lane << 26) | ic;
        qn += (uint32_t)__popcll(m);
        i += go ? 1u : 0u;
        go = go && i < v.n && v.pos[v.at(i)] < end;
        if (qn >= 64u) {
            wave_sync();
            drain(q + qn - 64u, 64u);
            qn -= 64u;
            wave_sync();
        }
    }
    if (qn) {
        wave_sync();
        drain(q, qn);
        wave_sync();
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

// ------------------------------------------------------------ count kernel

// A stratum whose window exceeds the LDS capacity: reads search the global
// records and count with global atomics (kept out of line so the LDS path's
// code is shared by both mates).
template <int BLOCK, bool CNA>
__device__ __forceinline__ void count_stratum_global(const job_dev& J, const seg_dev& S, const lds_tables& tab,
                                                  const stratum& s, uint32_t j, uint32_t c2, uint16_t* slots,
                                                  uint32_t tid, uint32_t* acc) {
    const glb_view v{J.pos + s.k0, J.lo + s.k0, J.hi + s.k0, J.info + s.k0, J.alt + S.out_off + s.k0,
                     J.depth + S.out_off + s.k0, s.n_win};
    for (uint32_t r = tid; r < (uint32_t)s.nj; r += BLOCK) {
        tmpl t = make_template<CNA>(J, S, tab, r, j, s.b, s.e);
        if (CNA) {
            if (!t.ok) continue;
            *acc += 1u;
        }
        for (uint32_t mate = 0; mate <= J.pe; ++mate)
            count_read(J, S, v, tab, t, mate ? t.x2 : t.x, r, j, c2, mate, slots, BLOCK);
    }
}

// CNA: some job has copy-number events (templates may be rejected and are
// counted); the plain instantiation carries none of that code.
template <int BLOCK, int CAP, bool CNA>
__global__ void __launch_bounds__(BLOCK) k_seq_count(const seq_launch L) {
    __shared__ uint32_t s_pos[CAP], s_lo[CAP], s_hi[CAP], s_info[CAP], s_alt[CAP], s_dep[CAP];
    __shared__ uint16_t s_bucket[NB_MAX];
    __shared__ uint16_t s_err[PSX_MAX_READ_ERRORS * BLOCK];
    __shared__ lds_tables s_tab;
    __shared__ uint32_t s_item, s_acc;
    __shared__ uint32_t s_queue[BLOCK / 64][128];  // per wave: (lane, record) entries
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    uint16_t* slots = s_err + tid;
    for (;;) {
        if (tid == 0) {
            s_item = (uint32_t)atomicAdd(L.queue, 1ull);
            s_acc = 0;
        }
        __syncthreads();
        const uint32_t item = __builtin_amdgcn_readfirstlane(s_item);
        if (item >= L.n_items) break;
        const item_dev it = uniform_at(L.items, item);
        const seg_dev S = uniform_at(L.segs, it.seg);
        const job_dev J = uniform_at(L.jobs, S.job);
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
                for (uint32_t base = wave * 64u; base < nj; base += BLOCK) {  // wave-uniform
                    const uint32_t r = base + lane;
                    bool valid = r < nj;
                    tmpl t = make_template<CNA>(J, S, s_tab, valid ? r : base, j, s.b, s.e);
                    if (CNA) {
                        valid = valid && t.ok;
                        acc += valid ? 1u : 0u;
                    }
                    for (uint32_t mate = 0; mate <= J.pe; ++mate)
                        count_mate_wave<BLOCK>(J, S, v, s_tab, t, valid, mate ? t.x2 : t.x, r, j, c2, mate, s_err,
                                               s_queue[wave], lane, wave);
                }
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
                count_stratum_global<BLOCK, CNA>(J, S, s_tab, s, j, c2, slots, tid, &acc);
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
__global__ void __launch_bounds__(256) k_seq_init(const job_dev* __restrict__ jobs, const pres_dev* __restrict__ ps,
                                                  uint32_t n_ps, uint64_t total,
                                                  unsigned long long* __restrict__ queue, uint32_t n_jobs) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q == 0) {
        *queue = 0ull;
        for (uint32_t k = 0; k < n_jobs; ++k)
            if (jobs[k].acc) *jobs[k].acc = 0ull;
    }
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
// SNV at each position and the read's errors applied; bases past the read
// end are 0. Oracle: the bases seq_read compares, materialised.
template <int NW, class V>
__device__ __forceinline__ void write_read(const job_dev& J, const V& v, const lds_tables& tab, uint32_t x,
                                           uint32_t leaf, uint32_t allele, uint32_t r, uint32_t j, uint32_t c2,
                                           uint32_t mate, uint16_t* slots, uint32_t stride, const uint32_t* s_ref,
                                           uint32_t ref_b, uint32_t* row) {
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
    // sets its base. Uniform trip count, branch-free read-modify-writes of the
    // lane's own row (an unchanged word when nothing applies).
    const uint32_t end = x + J.L;
    uint32_t i = v.first_u(x);
    bool go = i < v.n && v.pos[v.at(i)] < end;
    uint32_t last = 0xFFFFFFFFu;
    while (__ballot(go)) {
        const uint32_t ic = go ? i : 0u;
        const uint32_t p = v.pos[ic], in = v.info[ic];
        const bool apply = go && p != last && PSX_INFO_KIND(in) == PSX_SNV && PSX_INFO_ALLELE(in) == allele &&
                           carried(J, in, v.lo[ic], v.hi[ic], leaf);
        if (apply) {  // rare: only applying lanes touch their row (no bank conflicts from the others)
            const uint32_t o = p - x, s2 = (o & 15u) * 2u;
            row[2 + (o >> 4)] = (row[2 + (o >> 4)] & ~(3u << s2)) | (PSX_INFO_ALT(in) << s2);
        }
        last = apply ? p : last;
        i += go ? 1u : 0u;
        go = go && i < v.n && v.pos[v.at(i)] < end;
    }
    uint32_t ne = 0;
    if (J.err_n) ne = error_set_u(J, tab, r, j, c2, mate, slots, stride);
    for (uint32_t q = 0; q < ne; ++q) {  // rare: about one read in seven at e = 1e-3
        const uint32_t e = slots[q * stride];
        const uint32_t o = e & 0xFFFu, s2 = (o & 15u) * 2u;
        const uint32_t wv = row[2 + (o >> 4)];
        const uint32_t tb = (wv >> s2) & 3u, sub = e >> 12;
        const uint32_t nb = sub < tb ? sub : sub + 1;
        row[2 + (o >> 4)] = (wv & ~(3u << s2)) | (nb << s2);
    }
    if constexpr (NW == 0) row[0] = x;
    row[1] = leaf | (allele << 26) | (mate << 30) | ((ne > 0 ? 1u : 0u) << 31);
}

// STAGE: reads are built in LDS and copied out per pass (read_words small
// enough for 64 rows per wave); otherwise each lane builds its read in place
// in the output. NW > 0: reads of NW base words (rows of whole 16-byte
// vectors, built in registers); NW = 0: any read size.
template <int BLOCK, int CAP, bool STAGE, int NW>
__global__ void __launch_bounds__(BLOCK) k_seq_reads(const seq_launch L) {
    constexpr uint32_t WAVES = BLOCK / 64;
    __shared__ uint32_t s_pos[CAP], s_lo[CAP], s_hi[CAP], s_info[CAP];
    __shared__ uint16_t s_bucket[NB_MAX];
    __shared__ uint16_t s_err[PSX_MAX_READ_ERRORS * BLOCK];
    __shared__ uint32_t s_ref[REF_MAX];
    __shared__ lds_tables s_tab;
    __shared__ uint32_t s_item;
    extern __shared__ uint4 s_stage4[];  // STAGE: per wave 64 rows of read_words words
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    uint16_t* slots = s_err + tid;
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
            if (fits) stage_window<BLOCK, CAP>(J, s, s_pos, s_lo, s_hi, s_info, nullptr, nullptr, s_bucket, tid);
            else __syncthreads();
            const lds_view lv{s_pos, s_lo, s_hi, s_info, nullptr, nullptr, s_bucket, s.n_win, s.b};
            const glb_view gv{J.pos + s.k0, J.lo + s.k0, J.hi + s.k0, J.info + s.k0, nullptr, nullptr, s.n_win};
            const uint32_t nj = (uint32_t)s.nj;
            for (uint32_t c0 = wave * 64u; c0 < nj; c0 += WAVES * 64u) {
                const uint32_t rt = c0 + lane;
                uint32_t tx = 0, tx2 = 0, tla = 0;
                if (rt < nj) {
                    tmpl t = make_template(J, S, s_tab, rt, j, s.b, s.e);
                    tx = t.x;
                    tx2 = t.x2;
                    tla = leaf_of(J, S, t) | (t.allele << 26);
                }
                for (uint32_t p = 0; p < rpt; ++p) {
                    const uint32_t src = p * tpp + lane / rpt;
                    const uint32_t mate = lane & (rpt - 1u);
                    const uint32_t x1 = (uint32_t)__shfl((int)tx, (int)src);
                    const uint32_t x2 = (uint32_t)__shfl((int)tx2, (int)src);
                    const uint32_t la = (uint32_t)__shfl((int)tla, (int)src);
                    const uint32_t r = c0 + src;
                    const uint32_t t0 = c0 + p * tpp;  // first template of the pass
                    const uint32_t nt = t0 < nj ? min(tpp, nj - t0) : 0u;
                    const uint64_t first = S.read_base + (s.before + t0) * rpt;
                    uint32_t* row = STAGE ? stage + lane * RW : L.reads_out + (first + lane) * RW;
                    if (r < nj) {
                        const uint32_t x = mate ? x2 : x1;
                        if (fits)
                            write_read<NW>(J, lv, s_tab, x, la & 0x3FFFFFFu, la >> 26, r, j, c2, mate, slots,
                                           BLOCK, s_ref, ref_w0 << 4, row);
                        else
                            write_read<NW>(J, gv, s_tab, x, la & 0x3FFFFFFu, la >> 26, r, j, c2, mate, slots,
                                           BLOCK, s_ref, ref_w0 << 4, row);
                    }
                    if (STAGE) {
                        wave_sync();
                        const uint32_t words = nt * rpt * RW;
                        uint32_t* dst = L.reads_out + first * RW;
                        if (vec4) {
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

}  // namespace psxk
