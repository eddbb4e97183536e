// MutationEngine::place_mutations on gfx950 (src/mutation_engine.cpp:1085-1113).
//
//   k_edge_counts   one thread per forest node: Poisson SNV / indel counts
//   (scan)          exclusive sum of counts -> mutation ids per node
//   k_place         one thread per mutation: signature, SBS class, allele and
//                   a site whose trinucleotide matches the class (rejection on
//                   the 2-bit reference in HBM), or an indel
//   k_conflicts     infinite-sites model: after a footprint sort, a mutation
//                   touching an earlier-created one is redrawn (next attempt)
//   (sort + gather) SID order, sid_head flags, chromosome ranges
//
// Every draw is Philox(seed, counter) with the counter layout of the oracle,
// so the table is identical to the CPU restatement's.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "psx_id.h"
#include "psx_internal.h"

namespace {

constexpr uint32_t SITE_TRIES = 1u << 20;
constexpr uint32_t INDEL_TRIES = 1u << 20;
constexpr int IS_ROUNDS = 64;

enum : uint32_t { ST_OK = 0, ST_NO_SITE = 1, ST_NO_INDEL_ROOM = 2 };

struct gdev {
    const uint32_t* packed;
    const uint32_t* len;
    const uint64_t* woff;
    const uint64_t* boff;  // n_chr + 1
    uint64_t G;
    uint32_t n_chr;
};

__device__ __forceinline__ uint32_t gbase(const gdev& g, uint32_t c, uint32_t i) {
    return (g.packed[g.woff[c] + (i >> 4)] >> (2u * (i & 15u))) & 3u;
}

// chromosome of absolute base x: the last c with boff[c] <= x (binary search over n_chr + 1 offsets)
__device__ __forceinline__ void glocate(const gdev& g, uint64_t x, uint32_t& c, uint32_t& pos) {
    uint32_t lo = 0, hi = g.n_chr;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (g.boff[mid] <= x) lo = mid; else hi = mid;
    }
    c = lo;
    pos = (uint32_t)(x - g.boff[lo]);
}

// bases of one chromosome from HBM, for psxid::view
struct gbases {
    const uint32_t* words;  // the chromosome's first packed word
    __device__ __forceinline__ uint32_t operator()(uint32_t i) const { return (words[i >> 4] >> (2u * (i & 15u))) & 3u; }
};

// ------------------------------------------------------------ indel index
//
// One block per tile of ID_TILE positions of a chromosome: the tile's bases
// (with the halo the classes read: 5 before, 35 after) are staged in LDS, each
// thread classifies positions with psxid::view::sites_at, and the counts go
// through LDS counters to one global atomic per class and block. The list
// pass writes the sites of the listed classes, tagged with their class in
// bits 44.., for one radix sort.

constexpr uint32_t ID_TILE = 4096;
constexpr uint32_t ID_HALO = 64;  // bases staged before and after the tile (the classes reach 5 before, 35 after)
constexpr uint32_t ID_WORDS = (ID_TILE + 2 * ID_HALO) / 16 + 2;

struct id_args {
    const uint32_t* packed;
    const uint64_t* woff;
    const uint32_t* len;
    const uint64_t* boff;
    const uint32_t* tile_chr;
    const uint32_t* tile_first;
    unsigned long long* counts;   // [PSX_ID_CLASSES] (count pass)
    const uint32_t* listed;       // [PSX_ID_CLASSES] (list pass)
    unsigned long long* cursor;   // [PSX_ID_CLASSES] next slot of each listed class (list pass)
    uint64_t* list;
    uint32_t zero_listed;         // some class next to 0 copies is listed
};

struct lds_bases {
    const uint32_t* w;
    uint32_t w0;  // chromosome word of w[0]
    __device__ __forceinline__ uint32_t operator()(uint32_t i) const {
        return (w[(i >> 4) - w0] >> (2u * (i & 15u))) & 3u;
    }
};

template <bool LIST>
__global__ void __launch_bounds__(256) k_id_scan(const id_args a) {
    __shared__ uint32_t s_w[ID_WORDS];
    __shared__ uint32_t s_cnt[PSX_ID_CLASSES];
    const uint32_t c = a.tile_chr[blockIdx.x], first = a.tile_first[blockIdx.x];
    const uint32_t len = a.len[c];
    const uint32_t end = min(first + ID_TILE, len);
    const uint32_t lo = first > ID_HALO ? first - ID_HALO : 0u;
    const uint32_t hi = min(len, end + ID_HALO);
    const uint32_t w0 = lo >> 4, w1 = (hi + 15u) >> 4;
    const uint32_t* words = a.packed + a.woff[c];
    for (uint32_t i = threadIdx.x; i < w1 - w0; i += 256) s_w[i] = words[w0 + i];
    if (!LIST)
        for (uint32_t i = threadIdx.x; i < PSX_ID_CLASSES; i += 256) s_cnt[i] = 0;
    __syncthreads();
    const psxid::view<lds_bases> v{lds_bases{s_w, w0}, len};
    uint32_t z[6] = {0, 0, 0, 0, 0, 0};
    uint32_t cls[10];
    for (uint32_t q = first + threadIdx.x; q < end; q += 256) {
        const uint32_t n = v.sites_at(q, cls, z);
        if (!LIST) {
            for (uint32_t i = 0; i < n; ++i) atomicAdd(&s_cnt[cls[i]], 1u);
            continue;
        }
        const uint64_t gq = a.boff[c] + q;
        for (uint32_t i = 0; i < n; ++i) {
            if (!a.listed[cls[i]]) continue;
            const unsigned long long slot = atomicAdd(&a.cursor[cls[i]], 1ull);
            a.list[slot] = ((uint64_t)cls[i] << 44) | (gq << 10);
        }
        if (!a.zero_listed) continue;
        for (uint32_t zi = 0; zi < 6; ++zi) {
            const uint32_t k = psxid::zero_class(zi);
            if (!a.listed[k]) continue;
            const uint32_t np = psxid::payloads(k);
            for (uint32_t pl = 0; pl < np; ++pl) {
                if (!v.is_site(q, pl, k)) continue;
                const unsigned long long slot = atomicAdd(&a.cursor[k], 1ull);
                a.list[slot] = ((uint64_t)k << 44) | (gq << 10) | pl;
            }
        }
    }
    if (LIST) return;
    for (uint32_t zi = 0; zi < 6; ++zi)
        if (z[zi]) atomicAdd(&s_cnt[psxid::zero_class(zi)], z[zi]);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < PSX_ID_CLASSES; i += 256)
        if (s_cnt[i]) atomicAdd(&a.counts[i], (unsigned long long)s_cnt[i]);
}

__global__ void __launch_bounds__(256) k_id_untag(uint64_t* __restrict__ list, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i < n) list[i] &= (1ull << 44) - 1ull;
}

struct place_args {
    gdev g;
    uint64_t seed;
    uint32_t ploidy;
    uint32_t n_sbs;
    const uint64_t* expo_T;    // [n_epochs][n_sbs]
    const uint64_t* cls_T;     // [n_sbs][96]
    const uint64_t* len_T;
    uint32_t len_n;
    const uint32_t* node_epoch;
    const uint32_t* leaf_lo;
    const uint32_t* leaf_hi;
    uint32_t n_leaves;
    uint32_t n_nodes;
    // slot layout by id: drivers | germline | preneoplastic SNVs | preneoplastic indels | passengers
    uint64_t n_drv, n_germ, n_pre_snv, n_pre_indel;
    const uint32_t* drv_d;
    const uint32_t* drv_v;
    const psx_driver* drivers;
    const psx_germline* germ;
    uint32_t pre_sig;
    const uint32_t* cnt_snv;
    const uint64_t* node_off;
    uint64_t M;
    psx_mutation* recs;
    uint32_t* attempt;
    const uint8_t* redraw;  // nullptr = every id
    uint32_t* status;
    // indels from ID signatures (use_id): per-epoch exposure and per-signature class tables (classes without a
    // site in the reference have probability 0), the reference's sites per class and the listed classes' sites
    uint32_t use_id;
    uint32_t n_id;
    uint32_t pre_id;
    const uint64_t* id_expo_T;  // [n_epochs][n_id]
    const uint64_t* id_cls_T;   // [n_id][PSX_ID_CLASSES]
    const uint64_t* id_count;   // [PSX_ID_CLASSES]
    const uint32_t* id_listed;  // [PSX_ID_CLASSES]
    const uint64_t* id_off;     // [PSX_ID_CLASSES + 1]
    const uint64_t* id_list;
};

__device__ const uint32_t SUB_TARGET[6] = {0, 2, 3, 0, 1, 2};  // C>A C>G C>T T>A T>C T>G

__device__ uint32_t draw_snv(const place_args& a, uint32_t v, uint32_t k, uint32_t att, int fixed_sig,
                             psx_mutation& r) {
    const psx_u4 w = psx_draw(a.seed, k, v, att, PSX_P_SNV << 24);
    const uint32_t sig = fixed_sig >= 0 ? (uint32_t)fixed_sig
                                        : psx_cdf_pick(a.expo_T + (size_t)a.node_epoch[v] * a.n_sbs, a.n_sbs, w.x);
    const uint32_t cls = psx_cdf_pick(a.cls_T + (size_t)sig * 96u, 96u, w.y);
    const uint32_t allele = psx_uniform32(w.z, a.ploidy);
    const uint32_t sub = cls >> 4, five = (cls >> 2) & 3u, three = cls & 3u;
    const uint32_t mid = sub < 3 ? 1u : 3u, tgt = SUB_TARGET[sub];
    psx_u4 ws{0, 0, 0, 0};
    for (uint32_t t = 0; t < SITE_TRIES; ++t) {
        if ((t & 1u) == 0) ws = psx_draw(a.seed, k, v, att, (PSX_P_SITE << 24) | (t >> 1));
        const uint32_t hi = (t & 1u) ? ws.z : ws.x, lo = (t & 1u) ? ws.w : ws.y;
        uint32_t c, pos;
        glocate(a.g, psx_uniform64(hi, lo, a.g.G), c, pos);
        if (pos < 1 || pos + 1 >= a.g.len[c]) continue;
        const uint32_t m = gbase(a.g, c, pos);
        const uint32_t b5 = gbase(a.g, c, pos - 1), b3 = gbase(a.g, c, pos + 1);
        uint32_t alt;
        if (m == 1u || m == 3u) {
            if (!(b5 == five && m == mid && b3 == three)) continue;
            alt = tgt;
        } else {
            if (!(3u - b3 == five && 3u - m == mid && 3u - b5 == three)) continue;
            alt = 3u - tgt;
        }
        r.chr = c; r.pos = pos; r.kind = PSX_SNV; r.ref = (uint8_t)m; r.alt = (uint8_t)alt;
        r.len = 0; r.seq = 0; r.allele = (uint8_t)allele; r.cause = (uint16_t)sig;
        return ST_OK;
    }
    return ST_NO_SITE;
}

__device__ uint32_t draw_indel(const place_args& a, uint32_t v, uint32_t k, uint32_t att, psx_mutation& r) {
    for (uint32_t t = 0; t < INDEL_TRIES; ++t) {
        const psx_u4 w = psx_draw(a.seed, k, v, att, (PSX_P_INDEL << 24) | t);
        const uint32_t kind = (w.x >> 31) ? PSX_INS : PSX_DEL;
        const uint32_t allele = psx_uniform32(w.x << 1, a.ploidy);
        const uint32_t l = 1u + psx_cdf_pick(a.len_T, a.len_n, w.y);
        uint32_t c, pos;
        glocate(a.g, psx_uniform64(w.z, w.w, a.g.G), c, pos);
        if ((uint64_t)pos + l >= a.g.len[c]) continue;
        uint32_t seq = 0;
        if (kind == PSX_DEL) {
            for (uint32_t i = 0; i < l; ++i) seq |= gbase(a.g, c, pos + 1 + i) << (2u * i);
        } else {
            const psx_u4 s = psx_draw(a.seed, k, v, att, (PSX_P_INDEL << 24) | 0x400000u | t);
            seq = s.x & ((1u << (2u * l)) - 1u);
        }
        r.chr = c; r.pos = pos; r.kind = (uint8_t)kind; r.ref = (uint8_t)gbase(a.g, c, pos);
        r.alt = 0; r.len = (uint8_t)l; r.seq = seq; r.allele = (uint8_t)allele; r.cause = PSX_NO_CAUSE;
        return ST_OK;
    }
    return ST_NO_INDEL_ROOM;
}

// Indel from ID signatures: signature from the ID exposure of the node's
// epoch (or the preneoplastic one), class from the signature, allele uniform,
// site uniform among the class's sites: rejection over (position, payload)
// for dense classes, a uniform index into the class's sorted sites otherwise.
__device__ uint32_t draw_indel_id(const place_args& a, uint32_t v, uint32_t k, uint32_t att, int fixed_sig,
                                  psx_mutation& r) {
    const psx_u4 w = psx_draw(a.seed, k, v, att, PSX_P_IDCLS << 24);
    const uint32_t sig = fixed_sig >= 0 ? (uint32_t)fixed_sig
                                        : psx_cdf_pick(a.id_expo_T + (size_t)a.node_epoch[v] * a.n_id, a.n_id, w.x);
    const uint32_t cls = psx_cdf_pick(a.id_cls_T + (size_t)sig * PSX_ID_CLASSES, PSX_ID_CLASSES, w.y);
    const uint32_t allele = psx_uniform32(w.z, a.ploidy);
    uint32_t c = 0, q = 0, payload = 0;
    if (a.id_listed[cls]) {
        const psx_u4 ws = psx_draw(a.seed, k, v, att, PSX_P_IDSITE << 24);
        const uint64_t e = a.id_list[a.id_off[cls] + psx_uniform64(ws.x, ws.y, a.id_count[cls])];
        payload = (uint32_t)(e & 1023u);
        glocate(a.g, e >> 10, c, q);
    } else {
        const uint32_t np = psxid::payloads(cls);
        uint32_t t = 0;
        for (; t < SITE_TRIES; ++t) {
            const psx_u4 ws = psx_draw(a.seed, k, v, att, (PSX_P_IDSITE << 24) | t);
            glocate(a.g, psx_uniform64(ws.x, ws.y, a.g.G), c, q);
            payload = psx_uniform32(ws.z, np);
            const psxid::view<gbases> vw{gbases{a.g.packed + a.g.woff[c]}, a.g.len[c]};
            if (vw.is_site(q, payload, cls)) break;
        }
        if (t == SITE_TRIES) return ST_NO_SITE;
    }
    const psxid::cls_t kk = psxid::decode(cls);
    const psxid::view<gbases> vw{gbases{a.g.packed + a.g.woff[c]}, a.g.len[c]};
    r.chr = c; r.pos = q - 1; r.kind = kk.del ? PSX_DEL : PSX_INS; r.ref = (uint8_t)gbase(a.g, c, q - 1);
    r.alt = 0; r.len = (uint8_t)kk.u;
    r.seq = kk.zero ? psxid::zero_unit(kk.u, kk.b, payload) : vw.unit(q, kk.u);
    r.allele = (uint8_t)allele; r.cause = (uint16_t)(a.n_sbs + sig);
    return ST_OK;
}

__device__ uint32_t place_one(const place_args& a, uint64_t id, uint32_t att, psx_mutation& r) {
    r = psx_mutation{};
    uint64_t i = id;
    if (i < a.n_drv) {
        const uint32_t d = a.drv_d[i], v = a.drv_v[i];
        const psx_driver dr = a.drivers[d];
        uint32_t allele;
        if (dr.allele >= 0) allele = (uint32_t)dr.allele;
        else allele = psx_uniform32(psx_draw(a.seed, d, v, 0u, PSX_P_DRIVER << 24).x, a.ploidy);
        r.chr = dr.chr; r.pos = dr.pos; r.kind = PSX_SNV; r.ref = (uint8_t)gbase(a.g, dr.chr, dr.pos);
        r.alt = dr.alt; r.allele = (uint8_t)allele; r.cause = PSX_NO_CAUSE; r.nature = PSX_DRIVER;
        r.node = v; r.leaf_lo = a.leaf_lo[v]; r.leaf_hi = a.leaf_hi[v];
        return ST_OK;
    }
    i -= a.n_drv;
    if (i < a.n_germ) {
        const psx_germline gm = a.germ[i];
        r.chr = gm.chr; r.pos = gm.pos; r.kind = PSX_SNV; r.ref = (uint8_t)gbase(a.g, gm.chr, gm.pos);
        r.alt = gm.alt; r.allele = gm.allele; r.cause = PSX_NO_CAUSE; r.nature = PSX_GERMINAL;
        r.node = PSX_NO_NODE; r.leaf_lo = 0; r.leaf_hi = a.n_leaves;
        return ST_OK;
    }
    i -= a.n_germ;
    uint32_t st;
    if (i < a.n_pre_snv) {
        st = draw_snv(a, PSX_NO_NODE, (uint32_t)i, att, (int)a.pre_sig, r);
        r.nature = PSX_PRENEOPLASTIC; r.node = PSX_NO_NODE; r.leaf_lo = 0; r.leaf_hi = a.n_leaves;
        return st;
    }
    i -= a.n_pre_snv;
    if (i < a.n_pre_indel) {
        st = a.use_id ? draw_indel_id(a, PSX_NO_NODE, (uint32_t)i, att, (int)a.pre_id, r)
                      : draw_indel(a, PSX_NO_NODE, (uint32_t)i, att, r);
        r.nature = PSX_PRENEOPLASTIC; r.node = PSX_NO_NODE; r.leaf_lo = 0; r.leaf_hi = a.n_leaves;
        return st;
    }
    i -= a.n_pre_indel;
    // node of passenger i: last v with node_off[v] <= i
    uint32_t lo = 0, hi = a.n_nodes;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.node_off[mid] <= i) lo = mid + 1; else hi = mid;
    }
    const uint32_t v = lo - 1;
    const uint32_t q = (uint32_t)(i - a.node_off[v]);
    if (q < a.cnt_snv[v]) st = draw_snv(a, v, q, att, -1, r);
    else if (a.use_id) st = draw_indel_id(a, v, q - a.cnt_snv[v], att, -1, r);
    else st = draw_indel(a, v, q - a.cnt_snv[v], att, r);
    r.nature = PSX_PASSENGER; r.node = v; r.leaf_lo = a.leaf_lo[v]; r.leaf_hi = a.leaf_hi[v];
    return st;
}

__global__ void __launch_bounds__(256) k_place(const place_args a) {
    const uint64_t id = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= a.M) return;
    if (a.redraw && !a.redraw[id]) return;
    uint32_t att = 0;
    if (a.redraw) {
        att = a.attempt[id] + 1;
        a.attempt[id] = att;
    }
    psx_mutation r;
    const uint32_t st = place_one(a, id, att, r);
    if (st != ST_OK) atomicOr(a.status, st);
    a.recs[id] = r;
}

struct edge_args {
    uint64_t seed;
    uint32_t n_nodes;
    const uint32_t* parent;
    const uint32_t* species;
    const uint32_t* snv_k0;
    const uint32_t* snv_off;
    const uint32_t* snv_n;
    const uint32_t* ind_k0;
    const uint32_t* ind_off;
    const uint32_t* ind_n;
    const uint64_t* T;
    uint32_t* cnt_snv;
    uint64_t* cnt_tot;
};

// Passenger counts of the edge into node v (none for roots).
__global__ void __launch_bounds__(256) k_edge_counts(const edge_args a) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.n_nodes) return;
    if (a.parent[v] == PSX_NO_NODE) {
        a.cnt_snv[v] = 0;
        a.cnt_tot[v] = 0;
        return;
    }
    const psx_u4 w = psx_draw(a.seed, v, 0u, 0u, PSX_P_EDGE << 24);
    const uint32_t s = a.species[v];
    const uint32_t ns = a.snv_k0[s] + psx_cdf_pick(a.T + a.snv_off[s], a.snv_n[s], w.x);
    const uint32_t ni = a.ind_k0[s] + psx_cdf_pick(a.T + a.ind_off[s], a.ind_n[s], w.y);
    a.cnt_snv[v] = ns;
    a.cnt_tot[v] = (uint64_t)ns + ni;
}

__global__ void __launch_bounds__(256) k_footprint(const psx_mutation* __restrict__ recs, uint64_t M,
                                                   uint64_t* __restrict__ fkey, uint32_t* __restrict__ ids) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    fkey[i] = ((uint64_t)recs[i].chr << 32) | recs[i].pos;
    ids[i] = (uint32_t)i;
}

__device__ __forceinline__ uint32_t fp_end(const psx_mutation& r) {
    return r.pos + (r.kind == PSX_DEL ? (uint32_t)r.len : 0u);
}

// Footprints [pos, pos + deleted length] touching (gap < 2) an earlier-created
// mutation's footprint lose; drivers and germline mutations never lose.
__global__ void __launch_bounds__(256) k_conflicts(const psx_mutation* __restrict__ recs,
                                                   const uint64_t* __restrict__ fkey,
                                                   const uint32_t* __restrict__ ids, uint64_t M,
                                                   uint64_t n_fixed, uint8_t* __restrict__ loser,
                                                   uint32_t* __restrict__ n_lose) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    const uint32_t idj = ids[j];
    loser[idj] = 0;
    if (idj < n_fixed) return;
    const uint64_t kj = fkey[j];
    const uint32_t chr = (uint32_t)(kj >> 32), sj = (uint32_t)kj, ej = fp_end(recs[idj]);
    bool lose = false;
    for (uint64_t i = j; i-- > 0 && !lose;) {
        const uint64_t ki = fkey[i];
        const uint32_t si = (uint32_t)ki;
        if ((uint32_t)(ki >> 32) != chr || si + 12u < sj) break;
        const uint32_t idi = ids[i];
        if (idi < idj && (uint64_t)sj <= (uint64_t)fp_end(recs[idi]) + 1u && (uint64_t)si <= (uint64_t)ej + 1u)
            lose = true;
    }
    for (uint64_t i = j + 1; i < M && !lose; ++i) {
        const uint64_t ki = fkey[i];
        const uint32_t si = (uint32_t)ki;
        if ((uint32_t)(ki >> 32) != chr || (uint64_t)si > (uint64_t)ej + 1u) break;
        const uint32_t idi = ids[i];
        if (idi < idj && (uint64_t)sj <= (uint64_t)fp_end(recs[idi]) + 1u) lose = true;
    }
    if (lose) {
        loser[idj] = 1;
        atomicAdd(n_lose, 1u);
    }
}

__global__ void __launch_bounds__(256) k_sid_key(const psx_mutation* __restrict__ recs, uint64_t M,
                                                 uint64_t* __restrict__ key, uint32_t* __restrict__ ids) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const psx_mutation r = recs[i];
    const uint64_t payload = r.kind == PSX_SNV ? r.alt : (r.kind == PSX_INS ? r.seq : 0u);
    key[i] = ((uint64_t)r.chr << 58) | ((uint64_t)r.pos << 26) | ((uint64_t)r.kind << 24) |
             ((uint64_t)(r.len & 15u) << 20) | (payload & 0xFFFFFu);
    ids[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(256) k_gather(const psx_mutation* __restrict__ src,
                                                const uint64_t* __restrict__ key,
                                                const uint32_t* __restrict__ ids, uint64_t M,
                                                psx_mutation* __restrict__ dst) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    psx_mutation r = src[ids[j]];
    r.sid_head = (j == 0 || key[j] != key[j - 1]) ? 1 : 0;
    dst[j] = r;
}

__global__ void k_chr_begin(const uint64_t* __restrict__ key, uint64_t M, uint64_t* __restrict__ out) {
    const uint32_t c = threadIdx.x;
    if (c > PSX_MAX_CHROMOSOMES) return;
    const uint64_t target = c >= 64 ? ~0ull : ((uint64_t)c << 58);
    uint64_t lo = 0, hi = M;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (key[mid] < target) lo = mid + 1; else hi = mid;
    }
    out[c] = c >= 64 ? M : lo;
}

inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

const psx_id_index& psx_genome_id_index(psx_ctx* ctx, psx_genome* g) {
    std::lock_guard<std::mutex> lk(g->id_mu);
    if (g->id_index) return *g->id_index;
    auto ix = std::make_unique<psx_id_index>();
    hipStream_t s = ctx->stream;
    std::vector<uint32_t> tile_chr, tile_first;
    for (uint32_t c = 0; c < g->n_chr; ++c)
        for (uint64_t f = 0; f < g->len[c]; f += ID_TILE) {
            tile_chr.push_back(c);
            tile_first.push_back((uint32_t)f);
        }
    scratch_builder sb;
    const size_t o_len = sb.put_vec(g->len), o_woff = sb.put_vec(g->woff), o_boff = sb.put_vec(g->boff);
    const size_t o_tc = sb.put_vec(tile_chr), o_tf = sb.put_vec(tile_first);
    const std::vector<uint64_t> zeros(PSX_ID_CLASSES, 0);
    const size_t o_cnt = sb.put_vec(zeros);
    const size_t o_cur = sb.put_vec(zeros);
    const std::vector<uint32_t> zl(PSX_ID_CLASSES, 0);
    const size_t o_listed = sb.put_vec(zl);
    uint8_t* base = (uint8_t*)psx_scratch_upload(ctx, sb);
    id_args a{};
    a.packed = g->packed.as<uint32_t>();
    a.woff = (const uint64_t*)(base + o_woff);
    a.len = (const uint32_t*)(base + o_len);
    a.boff = (const uint64_t*)(base + o_boff);
    a.tile_chr = (const uint32_t*)(base + o_tc);
    a.tile_first = (const uint32_t*)(base + o_tf);
    a.counts = (unsigned long long*)(base + o_cnt);
    a.cursor = (unsigned long long*)(base + o_cur);
    a.listed = (const uint32_t*)(base + o_listed);
    const uint32_t n_tiles = (uint32_t)tile_chr.size();
    PSX_HIP(hipEventRecord(ctx->ev0, s));
    hipLaunchKernelGGL(k_id_scan<false>, dim3(n_tiles), dim3(256), 0, s, a);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipMemcpyAsync(ix->count, a.counts, sizeof ix->count, hipMemcpyDeviceToHost, s));
    PSX_HIP(hipStreamSynchronize(s));
    // a class is listed when a rejection search over (position, payload) would average more than DENSE_TRIES tries
    const uint64_t G = g->boff[g->n_chr];
    uint64_t total = 0;
    bool zero_listed = false;
    std::vector<uint64_t> cursor(PSX_ID_CLASSES);
    for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) {
        ix->off[k] = total;
        const unsigned __int128 need = (unsigned __int128)G * psxid::payloads(k);
        ix->listed[k] = ix->count[k] > 0 && (unsigned __int128)ix->count[k] * psxid::DENSE_TRIES < need;
        cursor[k] = total;
        if (ix->listed[k]) {
            total += ix->count[k];
            zero_listed |= psxid::decode(k).zero != 0;
        }
    }
    ix->off[PSX_ID_CLASSES] = total;
    ix->d_list.alloc(total * 8);
    if (total) {
        PSX_HIP(hipMemcpyAsync((void*)a.listed, ix->listed, sizeof ix->listed, hipMemcpyHostToDevice, s));
        PSX_HIP(hipMemcpyAsync(a.cursor, cursor.data(), 8 * cursor.size(), hipMemcpyHostToDevice, s));
        dbuf tagged;
        tagged.alloc(total * 8);
        a.list = tagged.as<uint64_t>();
        a.zero_listed = zero_listed ? 1u : 0u;
        hipLaunchKernelGGL(k_id_scan<true>, dim3(n_tiles), dim3(256), 0, s, a);
        PSX_HIP(hipGetLastError());
        size_t tmp_bytes = 0;
        PSX_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, tagged.as<uint64_t>(), ix->d_list.as<uint64_t>(),
                                                  (int)total, 0, 51, s));
        dbuf tmp;
        tmp.alloc(tmp_bytes);
        PSX_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, tagged.as<uint64_t>(), ix->d_list.as<uint64_t>(),
                                                  (int)total, 0, 51, s));
        hipLaunchKernelGGL(k_id_untag, dim3(grid_of(total)), dim3(256), 0, s, ix->d_list.as<uint64_t>(), total);
        PSX_HIP(hipGetLastError());
        std::vector<uint64_t> end(PSX_ID_CLASSES);
        PSX_HIP(hipMemcpyAsync(end.data(), a.cursor, 8 * end.size(), hipMemcpyDeviceToHost, s));
        PSX_HIP(hipStreamSynchronize(s));
        for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k)
            if (end[k] != ix->off[k + 1] && ix->listed[k])
                psx_throw(PSX_EHIP, "indel index: class %u listed %llu sites, counted %llu", k,
                          (unsigned long long)(end[k] - ix->off[k]), (unsigned long long)ix->count[k]);
    }
    PSX_HIP(hipEventRecord(ctx->ev1, s));
    PSX_HIP(hipStreamSynchronize(s));
    float ms = 0.f;
    PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ix->build_ms = ms;
    g->id_index = std::move(ix);
    return *g->id_index;
}

extern "C" psx_status psx_place_mutations(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                          const psx_engine_params* p, uint64_t seed, psx_mutations** out) {
    if (!ctx || !g || !f || !p || !out) return PSX_EINVAL;
    psx_mutations* mt = new psx_mutations();
    mt->ctx = ctx;
    psx_status st = psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        if (p->ploidy < 1 || p->ploidy > 255) psx_throw(PSX_EINVAL, "invalid ploidy");
        if (p->indel_max_len < 1 || p->indel_max_len > 10) psx_throw(PSX_EINVAL, "indel_max_len must be in [1, 10]");
        if (!(p->indel_len_p > 0.0 && p->indel_len_p <= 1.0)) psx_throw(PSX_EINVAL, "invalid indel length parameter");
        for (uint32_t v = 0; v < f->n_nodes; ++v)
            if (f->species[v] >= p->n_species) psx_throw(PSX_EINVAL, "node species %u has no rates", f->species[v]);
        bool need_snv = p->n_preneoplastic_snv > 0;
        for (uint32_t sp = 0; sp < p->n_species; ++sp) need_snv |= p->snv_rate[sp] > 0.0;
        if (need_snv && (p->n_sbs == 0 || p->n_epochs == 0))
            psx_throw(PSX_EINVAL, "SNVs need a signature and an exposure");
        if (p->n_preneoplastic_snv > 0 && p->preneoplastic_sbs >= p->n_sbs)
            psx_throw(PSX_EINVAL, "unknown preneoplastic signature");
        // indels from ID signatures: the genome's indel index (built once per genome) and the class tables
        bool use_id = false;
        const psx_id_index* idx = nullptr;
        if (p->n_id) {
            bool need_id = p->n_preneoplastic_indel > 0;
            for (uint32_t sp = 0; sp < p->n_species; ++sp) need_id |= p->indel_rate[sp] > 0.0;
            if (!p->id || (need_id && (!p->id_exposure || p->n_epochs == 0)))
                psx_throw(PSX_EINVAL, "indels need an ID signature and an exposure");
            if (p->n_preneoplastic_indel > 0 && p->preneoplastic_id >= p->n_id)
                psx_throw(PSX_EINVAL, "unknown preneoplastic ID signature");
            if (need_id) {
                use_id = true;
                idx = &psx_genome_id_index(ctx, const_cast<psx_genome*>(g));
            }
        }
        const uint64_t G = g->boff[g->n_chr];
        // ---- integer tables (host, from the double pmfs) -> one u64 array
        std::vector<uint64_t> T;
        auto append = [&](const uint64_t* src, size_t n) { size_t off = T.size(); T.insert(T.end(), src, src + n); return off; };
        std::vector<uint64_t> tmp;
        const size_t expo_off = T.size();
        for (uint32_t e = 0; e < p->n_epochs; ++e) {
            tmp.assign(p->n_sbs, 0);
            psx_cdf_table(p->exposure + (size_t)e * p->n_sbs, p->n_sbs, tmp.data());
            append(tmp.data(), tmp.size());
        }
        const size_t cls_off = T.size();
        for (uint32_t sg = 0; sg < p->n_sbs; ++sg) {
            tmp.assign(96, 0);
            psx_cdf_table(p->sbs + (size_t)sg * 96, 96, tmp.data());
            append(tmp.data(), tmp.size());
        }
        std::vector<double> lp(p->indel_max_len);
        for (uint32_t l = 1; l <= p->indel_max_len; ++l)
            lp[l - 1] = std::pow(1.0 - p->indel_len_p, (double)(l - 1)) * p->indel_len_p;
        tmp.assign(p->indel_max_len, 0);
        psx_cdf_table(lp.data(), p->indel_max_len, tmp.data());
        const size_t len_off = append(tmp.data(), tmp.size());
        size_t id_expo_off = 0, id_cls_off = 0;
        if (use_id) {
            id_expo_off = T.size();
            for (uint32_t e = 0; e < p->n_epochs; ++e) {
                tmp.assign(p->n_id, 0);
                psx_cdf_table(p->id_exposure + (size_t)e * p->n_id, p->n_id, tmp.data());
                append(tmp.data(), tmp.size());
            }
            id_cls_off = T.size();
            for (uint32_t sg = 0; sg < p->n_id; ++sg) {
                double pmf[PSX_ID_CLASSES];
                double mass = 0.0;
                for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) {
                    pmf[k] = idx->count[k] ? p->id[(size_t)sg * PSX_ID_CLASSES + k] : 0.0;
                    mass += pmf[k];
                }
                if (!(mass > 0.0)) psx_throw(PSX_EINVAL, "ID signature %u has no class with a site in the reference", sg);
                tmp.assign(PSX_ID_CLASSES, 0);
                psx_cdf_table(pmf, PSX_ID_CLASSES, tmp.data());
                append(tmp.data(), tmp.size());
            }
        }
        const double allelic = (double)((uint64_t)p->ploidy * G);
        std::vector<uint32_t> sk0(p->n_species), soff(p->n_species), sn(p->n_species);
        std::vector<uint32_t> ik0(p->n_species), ioff(p->n_species), in_(p->n_species);
        for (uint32_t sp = 0; sp < p->n_species; ++sp) {
            psx_itable a = psx_poisson_table(p->snv_rate[sp] * allelic);
            sk0[sp] = a.k0; sn[sp] = (uint32_t)a.T.size(); soff[sp] = (uint32_t)append(a.T.data(), a.T.size());
            psx_itable b = psx_poisson_table(p->indel_rate[sp] * allelic);
            ik0[sp] = b.k0; in_[sp] = (uint32_t)b.T.size(); ioff[sp] = (uint32_t)append(b.T.data(), b.T.size());
        }
        // ---- drivers, germline, node epochs (host)
        auto base_at = [&](uint32_t c, uint32_t i) -> uint32_t {
            uint32_t w = 0;
            PSX_HIP(hipMemcpy(&w, g->packed.as<uint32_t>() + g->woff[c] + i / 16, 4, hipMemcpyDeviceToHost));
            return (w >> (2 * (i % 16))) & 3u;
        };
        std::vector<uint32_t> drv_d, drv_v;
        for (uint32_t d = 0; d < p->n_drivers; ++d) {
            const psx_driver& dr = p->drivers[d];
            if (dr.chr >= g->n_chr || dr.pos >= g->len[dr.chr] || dr.alt > 3 ||
                (dr.allele >= 0 && (uint32_t)dr.allele >= p->ploidy) || dr.alt == base_at(dr.chr, dr.pos))
                psx_throw(PSX_EINVAL, "invalid driver %u", d);
            for (uint32_t v = 0; v < f->n_nodes; ++v) {
                const uint32_t par = f->parent[v];
                if (f->species[v] == dr.species && (par == PSX_NO_NODE || f->species[par] != dr.species)) {
                    drv_d.push_back(d);
                    drv_v.push_back(v);
                }
            }
        }
        for (uint32_t k = 0; k < p->n_germline; ++k) {
            const psx_germline& gm = p->germline[k];
            if (gm.chr >= g->n_chr || gm.pos >= g->len[gm.chr] || gm.alt > 3 || gm.allele >= p->ploidy ||
                gm.alt == base_at(gm.chr, gm.pos))
                psx_throw(PSX_EINVAL, "invalid germline mutation %u", k);
        }
        // copy-number drivers: on the edge into every node where the species first appears
        std::vector<psx_cna_event> cna;
        for (uint32_t k = 0; k < p->n_cna; ++k) {
            const psx_cna& c = p->cna[k];
            const bool wgd = c.kind == PSX_CNA_WGD;
            if (c.kind > PSX_CNA_WGD || c.species >= p->n_species ||
                (!wgd && (c.chr >= g->n_chr || c.len == 0 || c.start >= g->len[c.chr] ||
                          c.len > g->len[c.chr] - c.start ||
                          (c.allele >= p->ploidy && c.allele != PSX_CNA_RANDOM_ALLELE) ||
                          (c.kind == PSX_CNA_AMP && c.copies == 0))))
                psx_throw(PSX_EINVAL, "invalid copy-number driver %u", k);
            for (uint32_t v = 0; v < f->n_nodes; ++v) {
                const uint32_t par = f->parent[v];
                if (f->species[v] != c.species || (par != PSX_NO_NODE && f->species[par] == c.species)) continue;
                psx_cna_event e{};
                e.node = v;
                e.leaf_lo = f->leaf_lo[v];
                e.leaf_hi = f->leaf_hi[v];
                e.chr = wgd ? PSX_ALL_CHROMOSOMES : c.chr;
                e.start = wgd ? 0u : c.start;
                e.end = wgd ? 0u : c.start + c.len;
                e.allele = wgd ? 0 : c.allele;
                if (!wgd && c.allele == PSX_CNA_RANDOM_ALLELE)  // drawn per node: (driver, node, 1, DRIVER)
                    e.allele = (uint8_t)psx_uniform32(psx_draw(seed, k, v, 1u, PSX_P_DRIVER << 24).x, p->ploidy);
                e.kind = c.kind;
                e.copies = c.kind == PSX_CNA_AMP ? c.copies : 0;
                cna.push_back(e);
            }
        }
        // passenger CNAs (src/mutation_engine.cpp:441-487, 901-927): on the edge into every non-root node,
        // Poisson(cna_rate(species) x ploidy x G) events counted from word z of the node's edge draw (words x
        // and y count its SNVs and indels), each a uniform catalogue region on a uniform allele
        bool any_pcna = false;
        for (uint32_t sp = 0; sp < p->n_species && p->cna_rate; ++sp) {
            if (!(p->cna_rate[sp] >= 0.0) || !std::isfinite(p->cna_rate[sp]))
                psx_throw(PSX_EINVAL, "CNA rates must be non-negative");
            any_pcna |= p->cna_rate[sp] > 0.0;
        }
        if (any_pcna) {
            if (p->n_cna_catalogue == 0 || !p->cna_catalogue)
                psx_throw(PSX_EINVAL, "passenger CNAs need a CNA catalogue");
            for (uint32_t k = 0; k < p->n_cna_catalogue; ++k) {
                const psx_cna_region& r = p->cna_catalogue[k];
                if (r.chr >= g->n_chr || r.len == 0 || r.start >= g->len[r.chr] || r.len > g->len[r.chr] - r.start ||
                    (r.kind != PSX_CNA_AMP && r.kind != PSX_CNA_DEL))
                    psx_throw(PSX_EINVAL, "invalid passenger CNA catalogue region %u", k);
            }
            std::vector<psx_itable> ct(p->n_species);
            for (uint32_t sp = 0; sp < p->n_species; ++sp) ct[sp] = psx_poisson_table(p->cna_rate[sp] * allelic);
            for (uint32_t v = 0; v < f->n_nodes; ++v) {
                if (f->parent[v] == PSX_NO_NODE) continue;
                const psx_itable& t = ct[f->species[v]];
                if (t.T.size() == 1 && t.k0 == 0) continue;  // rate 0
                const psx_u4 w = psx_draw(seed, v, 0u, 0u, PSX_P_EDGE << 24);
                const uint32_t n_ev = t.k0 + psx_cdf_pick(t.T.data(), (uint32_t)t.T.size(), w.z);
                for (uint32_t k = 0; k < n_ev; ++k) {
                    const psx_u4 e = psx_draw(seed, k, v, 0u, PSX_P_PCNA << 24);
                    const psx_cna_region& r = p->cna_catalogue[psx_uniform32(e.x, p->n_cna_catalogue)];
                    psx_cna_event ev{};
                    ev.node = v;
                    ev.leaf_lo = f->leaf_lo[v];
                    ev.leaf_hi = f->leaf_hi[v];
                    ev.chr = r.chr;
                    ev.start = r.start;
                    ev.end = r.start + r.len;
                    ev.allele = (uint8_t)psx_uniform32(e.y, p->ploidy);
                    ev.kind = r.kind;
                    ev.copies = r.kind == PSX_CNA_AMP ? 1 : 0;
                    cna.push_back(ev);
                }
            }
        }
        // DFS order of the nodes; at one node the drivers in their order, then the passengers
        std::stable_sort(cna.begin(), cna.end(),
                         [](const psx_cna_event& a, const psx_cna_event& b) { return a.node < b.node; });
        psx_cna_check(cna, f->n_leaves, p->ploidy, g->len);
        mt->cna = std::move(cna);
        std::vector<uint32_t> epoch(f->n_nodes, 0);
        for (uint32_t v = 0; v < f->n_nodes; ++v)
            for (uint32_t e = 0; e < p->n_epochs; ++e)
                if (p->epoch_start[e] <= f->birth[v]) epoch[v] = e;
        // ---- descriptors into the scratch arena
        scratch_builder sb;
        const size_t o_T = sb.put_vec(T);
        const size_t o_len = sb.put_vec(g->len);
        const size_t o_woff = sb.put_vec(g->woff);
        const size_t o_boff = sb.put_vec(g->boff);
        const size_t o_sk0 = sb.put_vec(sk0), o_soff = sb.put_vec(soff), o_sn = sb.put_vec(sn);
        const size_t o_ik0 = sb.put_vec(ik0), o_ioff = sb.put_vec(ioff), o_in = sb.put_vec(in_);
        const size_t o_epoch = sb.put_vec(epoch);
        const size_t o_drvd = sb.put_vec(drv_d), o_drvv = sb.put_vec(drv_v);
        const size_t o_drivers = sb.put(p->drivers, sizeof(psx_driver) * p->n_drivers);
        const size_t o_germ = sb.put(p->germline, sizeof(psx_germline) * p->n_germline);
        const uint32_t zero = 0;
        const size_t o_status = sb.put(&zero, 4);
        size_t o_idc = 0, o_idl = 0, o_ido = 0;
        if (use_id) {
            o_idc = sb.put(idx->count, sizeof idx->count);
            o_idl = sb.put(idx->listed, sizeof idx->listed);
            o_ido = sb.put(idx->off, sizeof idx->off);
        }
        const size_t o_nlose = sb.put(&zero, 4);
        uint8_t* base = (uint8_t*)psx_scratch_upload(ctx, sb);
        auto at = [&](size_t off) { return (void*)(base + off); };

        PSX_HIP(hipEventRecord(ctx->ev0, s));
        const uint32_t n = f->n_nodes;
        dbuf d_cnt_snv, d_cnt_tot, d_off, d_tmp;
        d_cnt_snv.alloc((size_t)n * 4);
        d_cnt_tot.alloc((size_t)n * 8);
        d_off.alloc((size_t)n * 8);
        edge_args ea{};
        ea.seed = seed; ea.n_nodes = n; ea.parent = f->d_parent.as<uint32_t>(); ea.species = f->d_species.as<uint32_t>();
        ea.snv_k0 = (uint32_t*)at(o_sk0); ea.snv_off = (uint32_t*)at(o_soff); ea.snv_n = (uint32_t*)at(o_sn);
        ea.ind_k0 = (uint32_t*)at(o_ik0); ea.ind_off = (uint32_t*)at(o_ioff); ea.ind_n = (uint32_t*)at(o_in);
        ea.T = (uint64_t*)at(o_T); ea.cnt_snv = d_cnt_snv.as<uint32_t>(); ea.cnt_tot = d_cnt_tot.as<uint64_t>();
        hipLaunchKernelGGL(k_edge_counts, dim3(grid_of(n)), dim3(256), 0, s, ea);
        PSX_HIP(hipGetLastError());
        size_t tmp_bytes = 0;
        PSX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_cnt_tot.as<uint64_t>(), d_off.as<uint64_t>(),
                                                 (int)n, s));
        d_tmp.alloc(tmp_bytes);
        PSX_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp.p, tmp_bytes, d_cnt_tot.as<uint64_t>(), d_off.as<uint64_t>(),
                                                 (int)n, s));
        uint64_t last_off = 0, last_cnt = 0;
        PSX_HIP(hipMemcpyAsync(&last_off, d_off.as<uint64_t>() + (n - 1), 8, hipMemcpyDeviceToHost, s));
        PSX_HIP(hipMemcpyAsync(&last_cnt, d_cnt_tot.as<uint64_t>() + (n - 1), 8, hipMemcpyDeviceToHost, s));
        PSX_HIP(hipStreamSynchronize(s));
        const uint64_t n_pass = last_off + last_cnt;
        const uint64_t n_fixed = drv_d.size() + p->n_germline;
        const uint64_t M = n_fixed + p->n_preneoplastic_snv + p->n_preneoplastic_indel + n_pass;
        if (M >= 0xFFFFFFFFull) psx_throw(PSX_EINVAL, "too many mutations (%llu)", (unsigned long long)M);

        dbuf d_recs, d_attempt, d_redraw, d_key, d_key2, d_ids, d_ids2;
        d_recs.alloc(M * sizeof(psx_mutation));
        d_attempt.alloc(M * 4);
        d_redraw.alloc(M);
        d_key.alloc(M * 8); d_key2.alloc(M * 8);
        d_ids.alloc(M * 4); d_ids2.alloc(M * 4);
        PSX_HIP(hipMemsetAsync(d_attempt.p, 0, M * 4, s));
        place_args pa{};
        pa.g = gdev{g->packed.as<uint32_t>(), (uint32_t*)at(o_len), (uint64_t*)at(o_woff), (uint64_t*)at(o_boff), G,
                    g->n_chr};
        pa.seed = seed; pa.ploidy = p->ploidy; pa.n_sbs = p->n_sbs;
        pa.expo_T = (uint64_t*)at(o_T) + expo_off; pa.cls_T = (uint64_t*)at(o_T) + cls_off;
        pa.len_T = (uint64_t*)at(o_T) + len_off; pa.len_n = p->indel_max_len;
        pa.node_epoch = (uint32_t*)at(o_epoch);
        pa.leaf_lo = f->d_leaf_lo.as<uint32_t>(); pa.leaf_hi = f->d_leaf_hi.as<uint32_t>();
        pa.n_leaves = f->n_leaves; pa.n_nodes = n;
        pa.n_drv = drv_d.size(); pa.n_germ = p->n_germline;
        pa.n_pre_snv = p->n_preneoplastic_snv; pa.n_pre_indel = p->n_preneoplastic_indel;
        pa.drv_d = (uint32_t*)at(o_drvd); pa.drv_v = (uint32_t*)at(o_drvv);
        pa.drivers = (psx_driver*)at(o_drivers); pa.germ = (psx_germline*)at(o_germ);
        pa.pre_sig = p->preneoplastic_sbs;
        pa.cnt_snv = d_cnt_snv.as<uint32_t>(); pa.node_off = d_off.as<uint64_t>();
        pa.M = M; pa.recs = d_recs.as<psx_mutation>(); pa.attempt = d_attempt.as<uint32_t>();
        pa.redraw = nullptr; pa.status = (uint32_t*)at(o_status);
        if (use_id) {
            pa.use_id = 1; pa.n_id = p->n_id; pa.pre_id = p->preneoplastic_id;
            pa.id_expo_T = (uint64_t*)at(o_T) + id_expo_off; pa.id_cls_T = (uint64_t*)at(o_T) + id_cls_off;
            pa.id_count = (const uint64_t*)at(o_idc); pa.id_listed = (const uint32_t*)at(o_idl);
            pa.id_off = (const uint64_t*)at(o_ido); pa.id_list = idx->d_list.as<uint64_t>();
        }
        if (M) {
            hipLaunchKernelGGL(k_place, dim3(grid_of(M)), dim3(256), 0, s, pa);
            PSX_HIP(hipGetLastError());
        }
        auto check_status = [&] {
            uint32_t stv = 0;
            PSX_HIP(hipMemcpyAsync(&stv, pa.status, 4, hipMemcpyDeviceToHost, s));
            PSX_HIP(hipStreamSynchronize(s));
            if (stv & ST_NO_SITE) psx_throw(PSX_EINVAL, "no reference site matches a drawn SBS or ID class");
            if (stv & ST_NO_INDEL_ROOM) psx_throw(PSX_EINVAL, "no room for an indel");
        };
        check_status();
        size_t sort_bytes = 0, sort_bytes38 = 0;
        PSX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, d_key.as<uint64_t>(), d_key2.as<uint64_t>(),
                                                   d_ids.as<uint32_t>(), d_ids2.as<uint32_t>(), (int)M, 0, 64, s));
        PSX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes38, d_key.as<uint64_t>(), d_key2.as<uint64_t>(),
                                                   d_ids.as<uint32_t>(), d_ids2.as<uint32_t>(), (int)M, 0, 38, s));
        sort_bytes = std::max(sort_bytes, sort_bytes38);
        dbuf d_sort;
        d_sort.alloc(sort_bytes);
        if (p->infinite_sites && M > 1) {
            int round = 0;
            for (; round < IS_ROUNDS; ++round) {
                hipLaunchKernelGGL(k_footprint, dim3(grid_of(M)), dim3(256), 0, s, pa.recs, M, d_key.as<uint64_t>(),
                                   d_ids.as<uint32_t>());
                PSX_HIP(hipGetLastError());
                PSX_HIP(hipcub::DeviceRadixSort::SortPairs(d_sort.p, sort_bytes, d_key.as<uint64_t>(),
                                                           d_key2.as<uint64_t>(), d_ids.as<uint32_t>(),
                                                           d_ids2.as<uint32_t>(), (int)M, 0, 38, s));
                PSX_HIP(hipMemsetAsync(at(o_nlose), 0, 4, s));
                hipLaunchKernelGGL(k_conflicts, dim3(grid_of(M)), dim3(256), 0, s, pa.recs,
                                   (const uint64_t*)d_key2.p, (const uint32_t*)d_ids2.p, M, n_fixed,
                                   d_redraw.as<uint8_t>(), (uint32_t*)at(o_nlose));
                PSX_HIP(hipGetLastError());
                uint32_t n_lose = 0;
                PSX_HIP(hipMemcpyAsync(&n_lose, at(o_nlose), 4, hipMemcpyDeviceToHost, s));
                PSX_HIP(hipStreamSynchronize(s));
                if (n_lose == 0) break;
                place_args pr = pa;
                pr.redraw = d_redraw.as<uint8_t>();
                hipLaunchKernelGGL(k_place, dim3(grid_of(M)), dim3(256), 0, s, pr);
                PSX_HIP(hipGetLastError());
                check_status();
            }
            if (round == IS_ROUNDS) psx_throw(PSX_EINVAL, "infinite-sites placement did not converge");
        }
        // ---- SID order
        mt->n = M;
        mt->ploidy = p->ploidy;
        mt->n_leaves = f->n_leaves;
        mt->chr_len = g->len;
        mt->d_recs.alloc(M * sizeof(psx_mutation));
        std::vector<uint64_t> chr_begin(PSX_MAX_CHROMOSOMES + 1, M);
        if (M) {
            hipLaunchKernelGGL(k_sid_key, dim3(grid_of(M)), dim3(256), 0, s, pa.recs, M, d_key.as<uint64_t>(),
                               d_ids.as<uint32_t>());
            PSX_HIP(hipGetLastError());
            PSX_HIP(hipcub::DeviceRadixSort::SortPairs(d_sort.p, sort_bytes, d_key.as<uint64_t>(),
                                                       d_key2.as<uint64_t>(), d_ids.as<uint32_t>(),
                                                       d_ids2.as<uint32_t>(), (int)M, 0, 64, s));
            hipLaunchKernelGGL(k_gather, dim3(grid_of(M)), dim3(256), 0, s, pa.recs, (const uint64_t*)d_key2.p,
                               (const uint32_t*)d_ids2.p, M, mt->d_recs.as<psx_mutation>());
            PSX_HIP(hipGetLastError());
            dbuf d_cb;
            d_cb.alloc(8 * (PSX_MAX_CHROMOSOMES + 1));
            hipLaunchKernelGGL(k_chr_begin, dim3(1), dim3(128), 0, s, (const uint64_t*)d_key2.p, M,
                               d_cb.as<uint64_t>());
            PSX_HIP(hipGetLastError());
            PSX_HIP(hipMemcpyAsync(chr_begin.data(), d_cb.p, 8 * (PSX_MAX_CHROMOSOMES + 1), hipMemcpyDeviceToHost, s));
            PSX_HIP(hipStreamSynchronize(s));
        }
        PSX_HIP(hipEventRecord(ctx->ev1, s));
        PSX_HIP(hipStreamSynchronize(s));
        psx_mutations_finish(ctx, mt, chr_begin);
        PSX_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_place_ms = ms;
    });
    if (st != PSX_OK) { delete mt; return st; }
    *out = mt;
    return PSX_OK;
}
