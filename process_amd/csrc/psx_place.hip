// MutationEngine::place_mutations on gfx950 (src/mutation_engine.cpp:1085-1113).
//
//   k_edge_counts   one thread per forest node: Poisson SNV / indel counts
//   (scan)          exclusive sum of counts -> mutation ids per node
//   k_node_expand   the node of every passenger id
//   k_place         one thread per mutation: signature, SBS class, allele and
//                   a site drawn uniformly from the genome's context index
//                   list of the class's trinucleotide (one list read, driver
//                   loci redrawn), or an indel (ID classes: list or rejection)
//   k_conflicts     infinite-sites model, first round: after a sort by padded
//                   position, a mutation touching an earlier-created one
//                   loses; the losers are listed and redrawn (next attempt)
//   k_redraw        later rounds: the losers redrawn, each new footprint put
//                   into a hash table of moved mutations (by 32-base bucket)
//   k_check         only the redrawn footprints are checked, against the first
//                   round's sorted footprints and the hash table
//   (k_merge, k_sid_runs, k_gather) SID order: the first round's sorted
//                   footprints without the moved ones merged with the moved
//                   ones sorted; sid_head flags, chromosome ranges
//
// A call places a batch of forests (one for psx_place_mutations): each phase
// (node counts; placement and the first check; each later round; SID order)
// is launched for every job and ends in one synchronisation.
//
// Every draw is Philox(seed, counter) with the counter layout of the oracle,
// so the table is identical to the CPU restatement's.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "psx_id.h"
#include "psx_internal.h"

namespace {

constexpr uint32_t SITE_TRIES = 1u << 20;
constexpr uint32_t LIST_TRIES = 4096;  // draws of a context-index site before the class is given up as driver loci only
constexpr uint32_t INDEL_TRIES = 1u << 20;
constexpr int IS_ROUNDS = 256;  // checks of the infinite-sites model before placement gives up
constexpr uint32_t SPEC_ROUNDS = 4;  // infinite-sites rounds launched between two synchronisations

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
    uint32_t* attempt;       // attempts of every id (the first is 0)
    const uint32_t* node_of; // node of each passenger
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
    // SNV sites: the context index at the engine's sampling and the driver loci (global, ascending, unique)
    const uint32_t* sbs_list;
    uint32_t sbs_strand;      // list entries carry the site's strand in bit 31
    const uint64_t* sbs_kept;   // [32]
    const uint64_t* sbs_off;    // [33]
    const uint64_t* drv_pos;
    uint32_t n_drv_pos;
};

__device__ const uint32_t SUB_TARGET[6] = {0, 2, 3, 0, 1, 2};  // C>A C>G C>T T>A T>C T>G

__device__ __forceinline__ uint32_t node_epoch(const place_args& a, uint32_t v) {
    return a.node_epoch ? a.node_epoch[v] : 0u;
}

__device__ __forceinline__ bool driver_locus(const place_args& a, uint64_t gp) {
    uint32_t lo = 0, hi = a.n_drv_pos;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.drv_pos[mid] < gp) lo = mid + 1; else hi = mid;
    }
    return lo < a.n_drv_pos && a.drv_pos[lo] == gp;
}

// SNV: signature from the node's epoch (or the preneoplastic one), class from
// the signature (classes whose context has no indexed site have probability
// 0), allele uniform, site uniform over the class's context-index sites other
// than the driver loci; the strand of the site's base gives the alternative.
__device__ uint32_t draw_snv(const place_args& a, uint32_t v, uint32_t k, uint32_t att, int fixed_sig,
                             psx_mutation& r) {
    const psx_u4 w = psx_draw(a.seed, k, v, att, PSX_P_SNV << 24);
    const uint32_t sig = fixed_sig >= 0 ? (uint32_t)fixed_sig
                                        : psx_cdf_pick(a.expo_T + (size_t)node_epoch(a, v) * a.n_sbs, a.n_sbs, w.x);
    const uint32_t cls = psx_cdf_pick(a.cls_T + (size_t)sig * 96u, 96u, w.y);
    const uint32_t allele = psx_uniform32(w.z, a.ploidy);
    const uint32_t sub = cls >> 4, tgt = SUB_TARGET[sub];
    const uint32_t ctx = (sub >= 3u ? 16u : 0u) | (cls & 15u);
    const uint64_t n = a.sbs_kept[ctx];
    const uint32_t* list = a.sbs_list + a.sbs_off[ctx];
    for (uint32_t t = 0; t < LIST_TRIES && n; ++t) {
        const psx_u4 ws = psx_draw(a.seed, k, v, att, (PSX_P_SITE << 24) | t);
        const uint32_t en = list[psx_uniform64(ws.x, ws.y, n)];
        const uint32_t gp = a.sbs_strand ? en & 0x7FFFFFFFu : en;
        if (a.n_drv_pos && driver_locus(a, gp)) continue;
        uint32_t c, pos;
        glocate(a.g, gp, c, pos);
        // the site's base: C or G for the C contexts, T or A for the T ones (the strand bit tells which)
        const uint32_t m = a.sbs_strand ? (ctx < 16u ? ((en >> 31) ? 2u : 1u) : ((en >> 31) ? 0u : 3u))
                                        : gbase(a.g, c, pos);
        r.chr = c; r.pos = pos; r.kind = PSX_SNV; r.ref = (uint8_t)m;
        r.alt = (uint8_t)((m == 1u || m == 3u) ? tgt : 3u - tgt);
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
                                        : psx_cdf_pick(a.id_expo_T + (size_t)node_epoch(a, v) * a.n_id, a.n_id, w.x);
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
    const uint32_t v = a.node_of[i];
    const uint32_t q = (uint32_t)(i - a.node_off[v]);
    if (q < a.cnt_snv[v]) st = draw_snv(a, v, q, att, -1, r);
    else if (a.use_id) st = draw_indel_id(a, v, q - a.cnt_snv[v], att, -1, r);
    else st = draw_indel(a, v, q - a.cnt_snv[v], att, r);
    r.nature = PSX_PASSENGER; r.node = v; r.leaf_lo = a.leaf_lo[v]; r.leaf_hi = a.leaf_hi[v];
    return st;
}

// ------------------------------------------------------------- batched phases
//
// A call places a batch of jobs (forests over one genome) together. Every
// per-node and per-mutation array is one global array holding each job's
// range at its base, and each kernel runs once over every job: a thread finds
// its job by a binary search over the bases. Mutations are numbered globally
// (gid = the job's base + its id in the job). Sort keys carry the job index
// above the padded position (bit pbits up), so one radix sort orders every
// job's footprints (job-major, then by position), and since 2^pbits exceeds
// the padded genome by at least 16, footprints of different jobs are at
// least 16 apart and never touch.

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
    uint32_t* cnt_snv;  // at the job's node base
    uint64_t* cnt_tot;
};

// One job of the batch on the device.
struct fjob {
    place_args pa;        // recs, attempt, node_of, cnt_snv, node_off and status at the job's bases
    edge_args ea;         // its passenger counts
    uint64_t mbase;       // gid of its first mutation
    uint64_t n_fixed;     // drivers and germline mutations (never lose)
    uint32_t is;          // infinite-sites model
    uint32_t n_chr;
    psx_mutation* out;    // its table's records
    uint32_t *pos, *lo, *hi, *info;  // and the kernels' arrays of them
    uint64_t* chr_begin;  // its table's chromosome ranges (PSX_MAX_CHROMOSOMES + 1)
    uint32_t* bucket;     // its table's stratum index
};

struct batch {
    const fjob* jobs;
    uint32_t nj;
    uint32_t pbits;
    const uint64_t* nbase;     // node bases (nj + 1)
    const uint64_t* mbase;     // mutation bases (nj + 1)
    const uint64_t* pstart;    // padded chromosome starts: boff[c] + 16 c
    psx_mutation* recs;        // every job's records, at the mutation bases
};

// The job whose range holds g: the last j with base[j] <= g (base[nj] = total > g).
__device__ __forceinline__ uint32_t job_of(const uint64_t* __restrict__ base, uint32_t nj, uint64_t g) {
    uint32_t lo = 0, hi = nj;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= g) lo = mid; else hi = mid;
    }
    return lo;
}

// A job's descriptor read through the constant address space: with j
// wave-uniform its fields come by scalar loads into SGPRs.
__device__ __forceinline__ const fjob& cjob(const batch& B, uint32_t j) {
    return *(const fjob*)((const __attribute__((address_space(4))) fjob*)B.jobs + j);
}

__device__ __forceinline__ uint64_t first_lane(uint64_t x) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// job_of for lanes holding increasing g (consecutive ids of a wave): the
// first active lane's job by a scalar search, lanes past that job's end (a
// wave at a job boundary) by their own.
__device__ __forceinline__ uint32_t job_of_seq(const uint64_t* __restrict__ base, uint32_t nj, uint64_t g) {
    const __attribute__((address_space(4))) uint64_t* cb = (const __attribute__((address_space(4))) uint64_t*)base;
    const uint64_t g0 = first_lane(g);
    uint32_t lo = 0, hi = nj;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cb[mid] <= g0) lo = mid; else hi = mid;
    }
    return g < cb[lo + 1] ? lo : job_of(base, nj, g);
}

// f(j) with j wave-uniform: one pass per distinct job among the wave's
// active lanes (one unless the wave straddles a job boundary), each lane in
// the pass of its own job. The loop runs on a wave-uniform mask, so the
// passes stay in the loop with the job index in an SGPR.
template <class F>
__device__ __forceinline__ void per_job(uint32_t j, F&& f) {
    uint64_t left = __ballot(1);
    while (left) {
        const uint32_t ju = (uint32_t)__builtin_amdgcn_readlane((int)j, (int)__builtin_ctzll(left));
        uint32_t jv = j;
        asm volatile("" : "+v"(jv));  // an opaque copy made after ju: where the two are equal the
        const bool mine = ju == jv;   // compiler substitutes ju for it, not the divergent j for ju
        if (mine) f(ju);
        left &= ~__ballot(mine);
    }
}

// Sort key of a record of job j: the job above the padded position (positions
// of different chromosomes at least 16 apart, so footprints of at most 11
// bases on different chromosomes never touch).
__device__ __forceinline__ uint64_t jkey(const batch& B, uint32_t j, const psx_mutation& r) {
    return ((uint64_t)j << B.pbits) + B.pstart[r.chr] + r.pos;
}

// Passenger counts of the edge into every node of every job (none for roots).
__global__ void __launch_bounds__(256) k_edge_counts(const batch B, uint64_t n) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    per_job(job_of_seq(B.nbase, B.nj, g), [&](uint32_t j) {
        const edge_args& a = cjob(B, j).ea;
        const uint32_t v = (uint32_t)(g - B.nbase[j]);
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
    });
}

// After an exclusive scan of every node's count (off, n + 1 entries): each
// node's first passenger within its job, and each job's passenger count.
__global__ void __launch_bounds__(256) k_node_bases(const batch B, uint64_t n, const uint64_t* __restrict__ off,
                                                    uint64_t* __restrict__ off_job, uint64_t* __restrict__ pcount) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) off_job[g] = off[g] - off[B.nbase[job_of_seq(B.nbase, B.nj, g)]];
    if (g < B.nj) pcount[g] = off[B.nbase[g + 1]] - off[B.nbase[g]];
}

// node_of[off[g] + k] = v (k < cnt[g]) for node v of job j (g = nbase[j] + v),
// node_of being every job's passenger array (each at its passenger base
// off[nbase[j]]). A wave takes 64 nodes and writes their passengers'
// entries together, 64 consecutive entries a pass: each lane finds its
// entry's node among the 64 first offsets in LDS, so nodes with many
// passengers need no long loop of one lane.
__global__ void __launch_bounds__(256) k_node_expand(const batch B, uint64_t n, const uint64_t* __restrict__ off,
                                                     uint32_t* __restrict__ node_of) {
    __shared__ uint64_t s_off[4][64];
    __shared__ uint32_t s_v[4][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const uint64_t g0 = ((uint64_t)blockIdx.x * blockDim.x) + wv * 64u;
    if (g0 >= n) return;
    const uint64_t g = g0 + lane;
    const uint64_t gl = g < n ? g : n - 1u;
    s_off[wv][lane] = g < n ? off[g] : off[n];
    s_v[wv][lane] = (uint32_t)(gl - B.nbase[job_of_seq(B.nbase, B.nj, gl)]);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const uint64_t start = s_off[wv][0];
    const uint64_t end = g0 + 64u < n ? off[g0 + 64u] : off[n];
    for (uint64_t i = start + lane; i < end; i += 64u) {
        uint32_t lo = 0;  // last node of the 64 whose first entry is at or before i
#pragma unroll
        for (uint32_t step = 32; step; step >>= 1)
            if (s_off[wv][lo + step] <= i) lo += step;
        node_of[i] = s_v[wv][lo];
    }
}

// Sort entries of the infinite-sites check: key = jkey, value = id in the job
// | deleted length << 28 (ids are below 2^28).
constexpr uint32_t ID_BITS = 28, ID_MASK = (1u << ID_BITS) - 1u, NO_ENTRY = 0xFFFFFFFFu;

// deleted length: a footprint covers [start, start + length] (at most 11 bases)
__device__ __forceinline__ uint32_t fp_len(const psx_mutation& r) {
    return r.kind == PSX_DEL ? (uint32_t)(r.len & 15u) : 0u;
}

__device__ __forceinline__ uint32_t fp_val(uint32_t id, const psx_mutation& r) { return id | (fp_len(r) << ID_BITS); }

// footprints [sa, sa + la] and [sb, sb + lb] touch: gap below 2 bases
__device__ __forceinline__ bool fp_touch(uint64_t sa, uint32_t la, uint64_t sb, uint32_t lb) {
    return sb <= sa + la + 1u && sa <= sb + lb + 1u;
}

// One thread per mutation of every job, first attempt, and its sort entry.
__global__ void __launch_bounds__(256) k_place(const batch B, uint64_t n, uint64_t* __restrict__ fkey,
                                               uint32_t* __restrict__ fval) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    per_job(job_of_seq(B.mbase, B.nj, g), [&](uint32_t j) {
        const place_args& a = cjob(B, j).pa;
        const uint64_t id = g - B.mbase[j];
        psx_mutation r;
        const uint32_t st = place_one(a, id, 0u, r);
        if (st != ST_OK) atomicOr(a.status, st);
        a.recs[id] = r;
        fkey[g] = jkey(B, j, r);
        fval[g] = fp_val((uint32_t)id, r);
    });
}

// A loser of check number `stamp` (1 for the first): lose[gid] holds the
// number of the last check it lost (cleared once per call), so every loser is
// listed once per check without clearing flags between checks. Beside each
// listed gid: the index of its first-check entry when it has not moved yet
// (found through that entry), else NO_ENTRY.
__device__ __forceinline__ void mark_loser(uint32_t* lose, uint32_t stamp, uint32_t* list, uint32_t* apos,
                                           uint32_t* n_lose, uint32_t gid, uint32_t ap) {
    if (atomicExch(&lose[gid], stamp) != stamp) {
        const uint32_t slot = atomicAdd(n_lose, 1u);
        list[slot] = gid;
        apos[slot] = ap;
    }
}

// Appending kernels take APPEND_ITEMS items per thread (item k of thread t of
// block b: b * 256 * APPEND_ITEMS + t + 256 k) and reserve their output slots
// with one atomic per block: atomics on one address serialise (about 12 ns
// each), so one per item or per wave would cost more than the scan itself.
constexpr uint32_t APPEND_ITEMS = 16;

// The calling thread's first slot for n items (every thread of the block calls it).
__device__ __forceinline__ uint32_t block_reserve(uint32_t n, uint32_t* counter) {
    using scan = hipcub::BlockScan<uint32_t, 256>;
    __shared__ typename scan::TempStorage tmp;
    __shared__ uint32_t s_base;
    uint32_t excl = 0, total = 0;
    scan(tmp).ExclusiveSum(n, excl, total);
    if (threadIdx.x == 0) s_base = total ? atomicAdd(counter, total) : 0u;
    __syncthreads();
    return s_base + excl;
}

// First round of the infinite-sites model over every job's footprints sorted
// by key: footprints touching (gap < 2) an earlier-created mutation's
// footprint of the same job lose (neighbours of other jobs are too far to
// touch); drivers and germline mutations (ids below n_fixed) never lose, nor
// does anything of a job without the model. The losers' stamps are set to 1
// (the caller clears them all first) and their gids appended to the redraw
// list with their entry index.
__device__ __forceinline__ bool loses(const batch& B, const uint64_t* __restrict__ fkey,
                                      const uint32_t* __restrict__ fval, uint64_t n, uint64_t t, uint32_t& gid) {
    const uint64_t sj = fkey[t];
    const fjob& J = B.jobs[(uint32_t)(sj >> B.pbits)];
    const uint32_t vj = fval[t], lj = vj >> ID_BITS, idj = vj & ID_MASK;
    gid = (uint32_t)(J.mbase + idj);
    if (!J.is || idj < J.n_fixed) return false;
    const uint64_t ej = sj + lj;
    for (uint64_t i = t; i-- > 0;) {
        const uint64_t si = fkey[i];
        if (si + 12u < sj) break;
        const uint32_t vi = fval[i];
        if ((vi & ID_MASK) < idj && fp_touch(si, vi >> ID_BITS, sj, lj)) return true;
    }
    for (uint64_t i = t + 1; i < n; ++i) {
        const uint64_t si = fkey[i];
        if (si > ej + 1u) break;
        const uint32_t vi = fval[i];
        if ((vi & ID_MASK) < idj && fp_touch(si, vi >> ID_BITS, sj, lj)) return true;
    }
    return false;
}

__global__ void __launch_bounds__(256) k_conflicts(const batch B, const uint64_t* __restrict__ fkey,
                                                   const uint32_t* __restrict__ fval, uint64_t n,
                                                   uint32_t* __restrict__ lose, uint32_t* __restrict__ list,
                                                   uint32_t* __restrict__ apos, uint32_t* __restrict__ n_lose) {
    const uint64_t b = (uint64_t)blockIdx.x * 256u * APPEND_ITEMS + threadIdx.x;
    uint32_t mask = 0, c = 0, gids[APPEND_ITEMS];
#pragma unroll
    for (uint32_t k = 0; k < APPEND_ITEMS; ++k) {
        const uint64_t t = b + 256u * k;
        gids[k] = 0;
        if (t < n && loses(B, fkey, fval, n, t, gids[k])) {
            mask |= 1u << k;
            ++c;
        }
    }
    uint32_t slot = block_reserve(c, n_lose);
#pragma unroll
    for (uint32_t k = 0; k < APPEND_ITEMS; ++k)
        if (mask & (1u << k)) {
            lose[gids[k]] = 1u;
            list[slot] = gids[k];
            apos[slot] = (uint32_t)(b + 256u * k);
            ++slot;
        }
}

__device__ __forceinline__ uint64_t lower_bound_u64(const uint64_t* a, uint64_t n, uint64_t x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t upper_bound_u64(const uint64_t* a, uint64_t n, uint64_t x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] <= x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Moved mutations by the 32-base bucket of their key, in an open-addressing
// table with linear probing (slots hold gids, NO_ENTRY when empty; the host
// keeps it at most half full). Slots are never cleared while rounds run: a
// mutation that moves again leaves its old entry behind, which no longer
// touches anything because footprints are always read from the mutation's
// current record.
constexpr uint32_t BUCKET_SHIFT = 5;

__device__ __forceinline__ uint32_t tab_home(uint64_t bucket, uint32_t mask) {
    return (uint32_t)((bucket * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

__device__ __forceinline__ void tab_insert(uint32_t* tab, uint32_t mask, uint64_t key, uint32_t gid) {
    uint32_t h = tab_home(key >> BUCKET_SHIFT, mask);
    for (uint32_t k = 0; k <= mask; ++k, h = (h + 1u) & mask)
        if (atomicCAS(&tab[h], NO_ENTRY, gid) == NO_ENTRY) return;
}

// Device state of the infinite-sites rounds of a batch: rounds are launched
// several at a time without the host in between, so each reads its loser
// count here.
struct round_io {
    uint32_t n_moved;   // mutations moved so far (entries of the moved list)
    uint32_t tab_used;  // table entries since it was last cleared
    uint32_t halt;      // first round that found the table too full to run (0: none)
    uint32_t pad;
    uint32_t n_lose[IS_ROUNDS + 1];  // losers of check r (0: the first check)
};

struct redraw_args {
    uint32_t round;          // r >= 1: redraws the losers of check r - 1
    const uint32_t* list;    // losers of check r - 1 (gids)
    const uint32_t* lapos;   // and their first-check entries (NO_ENTRY once moved)
    uint32_t* moved;         // set for every redrawn gid
    uint32_t* mlist;         // gids in the order they first moved
    uint32_t* mapos;         // and the index of each one's first-check entry
    const uint64_t* akey;    // the first check's sorted entries: a mutation's entry is made NO_ENTRY
    uint32_t* aval;          // when it first moves
    uint64_t n;
    round_io* io;
    uint32_t* tab;
    uint32_t tab_mask;
};

// Round r: the losers' next attempts (grid-stride), each new footprint into
// the table. The round does nothing after a halt, and halts when its entries
// would fill the table beyond half.
__global__ void __launch_bounds__(256) k_redraw(const batch B, const redraw_args b) {
    round_io* io = b.io;
    const uint32_t n = io->n_lose[b.round - 1];
    if (io->halt || n == 0) return;
    if ((uint64_t)io->tab_used + n > (uint64_t)(b.tab_mask + 1u) / 2u) {
        if (blockIdx.x == 0 && threadIdx.x == 0) io->halt = b.round;
        return;
    }
    for (uint32_t base = blockIdx.x * 256u; base < n; base += gridDim.x * 256u) {
        const uint32_t t = base + threadIdx.x;
        const bool on = t < n;
        const uint32_t gid = on ? b.list[t] : 0u;
        const bool first = on && atomicExch(&b.moved[gid], 1u) == 0u;
        const uint32_t slot = block_reserve(first ? 1u : 0u, &io->n_moved);
        const uint32_t j = job_of(B.mbase, B.nj, gid);
        if (first) {
            uint64_t e = b.lapos[t];
            if (e == NO_ENTRY) {  // not listed with it: found from the first-check footprint, still in its record
                e = lower_bound_u64(b.akey, b.n, jkey(B, j, B.recs[gid]));
                const uint32_t id = (uint32_t)(gid - B.mbase[j]);
                while (e + 1 < b.n && (b.aval[e] & ID_MASK) != id) ++e;  // among the entries at that key
            }
            b.aval[e] = NO_ENTRY;
            b.mlist[slot] = gid;
            b.mapos[slot] = (uint32_t)e;
        }
        if (on) {
            const place_args& a = B.jobs[j].pa;
            const uint64_t id = gid - B.mbase[j];
            const uint32_t att = a.attempt[id] + 1u;
            a.attempt[id] = att;
            psx_mutation r;
            const uint32_t st = place_one(a, id, att, r);
            if (st != ST_OK) atomicOr(a.status, st);
            a.recs[id] = r;
            tab_insert(b.tab, b.tab_mask, jkey(B, j, r), gid);
        }
        __syncthreads();  // block_reserve's shared slot is reused by the next pass
    }
}

// The table rebuilt (after the host cleared it) from the moved mutations'
// current footprints; the halt, if any, is lifted.
__global__ void __launch_bounds__(256) k_tab_fill(const batch B, const uint32_t* __restrict__ mlist, uint32_t n,
                                                  uint32_t* tab, uint32_t mask, round_io* io) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        io->tab_used = n;
        io->halt = 0;
    }
    if (t >= n) return;
    const uint32_t gid = mlist[t];
    tab_insert(tab, mask, jkey(B, job_of(B.mbase, B.nj, gid), B.recs[gid]), gid);
}

struct check_args {
    const uint64_t* akey;  // first-round entries, sorted (entries of moved mutations are NO_ENTRY)
    const uint32_t* aval;
    uint64_t n;
    const uint32_t* tab;
    uint32_t tab_mask;
    uint32_t round;        // r >= 1: checks the footprints redrawn in round r (check r)
    const uint32_t* list;  // losers of check r - 1, redrawn in round r
    round_io* io;
    uint32_t* lose;
    uint32_t* out;         // losers of check r
    uint32_t* out_apos;    // and their first-check entries
};

// Later rounds: a pair of touching footprints has a member redrawn in the
// last redraw (a touching pair left unchanged would have made its later
// member lose and be redrawn), so only those are checked: each against the
// first-round entries of unmoved mutations and the moved ones in the table
// (touching entries are of the same job, whose ids compare as gids do). The
// later member of every touching pair loses, as in a full pass.
__device__ __forceinline__ void check_one(const batch& B, const check_args& a, uint32_t gc) {
    const uint32_t stamp = a.round + 1u;
    uint32_t* n_out = &a.io->n_lose[a.round];
    const uint32_t j = job_of(B.mbase, B.nj, gc);
    const uint64_t mb = B.mbase[j];
    const uint32_t idc = (uint32_t)(gc - mb);
    const psx_mutation rc = B.recs[gc];
    const uint64_t sc = jkey(B, j, rc);
    const uint32_t lc = fp_len(rc);
    const uint64_t from = sc >= 11u ? sc - 11u : 0u, to = sc + lc + 1u;  // starts of touching footprints
    bool l = false;
    for (uint64_t i = lower_bound_u64(a.akey, a.n, from); i < a.n; ++i) {
        const uint64_t si = a.akey[i];
        if (si > to) break;
        const uint32_t vi = a.aval[i], idi = vi & ID_MASK;
        if (vi == NO_ENTRY || !fp_touch(si, vi >> ID_BITS, sc, lc)) continue;
        if (idi < idc) l = true;
        else mark_loser(a.lose, stamp, a.out, a.out_apos, n_out, (uint32_t)(mb + idi), (uint32_t)i);
    }
    for (uint64_t bk = from >> BUCKET_SHIFT; bk <= to >> BUCKET_SHIFT; ++bk) {
        uint32_t h = tab_home(bk, a.tab_mask);
        for (uint32_t k = 0; k <= a.tab_mask; ++k, h = (h + 1u) & a.tab_mask) {
            const uint32_t gi = a.tab[h];
            if (gi == NO_ENTRY) break;
            if (gi == gc) continue;
            const psx_mutation ri = B.recs[gi];
            if (!fp_touch(jkey(B, job_of(B.mbase, B.nj, gi), ri), fp_len(ri), sc, lc)) continue;
            if (gi < gc) l = true;
            else mark_loser(a.lose, stamp, a.out, a.out_apos, n_out, gi, NO_ENTRY);
        }
    }
    if (l) mark_loser(a.lose, stamp, a.out, a.out_apos, n_out, gc, NO_ENTRY);
}

// Check r (grid-stride over the footprints redrawn in round r); nothing after a halt.
__global__ void __launch_bounds__(256) k_check(const batch B, const check_args a) {
    round_io* io = a.io;
    const uint32_t n = io->n_lose[a.round - 1];
    if (io->halt || n == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) io->tab_used += n;  // k_redraw's entries
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        check_one(B, a, a.list[t]);
}

// SID order: the moved mutations' current keys (sorted next) ...
__global__ void __launch_bounds__(256) k_moved_keys(const batch B, const uint32_t* __restrict__ mlist, uint32_t n,
                                                    uint64_t* __restrict__ key, uint32_t* __restrict__ gids) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint32_t gid = mlist[t];
    key[t] = jkey(B, job_of(B.mbase, B.nj, gid), B.recs[gid]);
    gids[t] = gid;
}

// ... then the output slots by one scan over w[0..n]: w[i] = 1 for a
// first-check entry still current (0 for the moved ones') plus the number of
// moved keys whose upper bound among the first-check keys is i. With S the
// exclusive sum of w, a current entry i goes to S[i + 1] - 1 and the k-th
// moved key (upper bound u, the first of its bound at k0) to S[u] - k0 + k:
// both count the current entries and moved keys before them, ties put
// first-check entries first. The output is job-major, each job's range at its
// mutation base.
__global__ void __launch_bounds__(256) k_merge_marks(const uint32_t* __restrict__ mapos,
                                                     const uint64_t* __restrict__ akey, uint64_t n,
                                                     const uint64_t* __restrict__ bkey, uint32_t nb, uint32_t* w,
                                                     uint32_t* __restrict__ ub) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    atomicSub(&w[mapos[k]], 1u);
    const uint64_t u = upper_bound_u64(akey, n, bkey[k]);
    ub[k] = (uint32_t)u;
    atomicAdd(&w[u], 1u);
}

struct merge_args {
    const uint64_t* akey;  // first-check entries, sorted (moved ones NO_ENTRY)
    const uint32_t* aval;
    uint64_t n;
    const uint32_t* S;     // exclusive sum of w (n + 1 entries)
    const uint64_t* bkey;  // moved mutations' current keys, sorted
    const uint32_t* bval;  // and gids
    const uint32_t* ub;    // their upper bounds among akey
    uint32_t nb;
    uint64_t* okey;
    uint32_t* ogid;
};

// One thread per entry of either; okey/ogid receive n entries. Without moved
// mutations the first check's order is the output.
__global__ void __launch_bounds__(256) k_merge(const batch B, const merge_args a) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < a.n) {
        const uint32_t v = a.aval[t];
        if (v == NO_ENTRY) return;
        const uint64_t k = a.akey[t];
        const uint64_t o = a.nb ? a.S[t + 1] - 1u : t;
        a.okey[o] = k;
        a.ogid[o] = (uint32_t)(B.jobs[(uint32_t)(k >> B.pbits)].mbase + (v & ID_MASK));
    } else if (t < a.n + a.nb) {
        const uint32_t k = (uint32_t)(t - a.n), u = a.ub[k];
        uint32_t lo = 0, hi = k;  // first moved key with the same bound
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (a.ub[mid] < u) lo = mid + 1; else hi = mid;
        }
        const uint64_t o = (uint64_t)a.S[u] - lo + k;
        a.okey[o] = a.bkey[k];
        a.ogid[o] = a.bval[k];
    }
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, uint64_t n, uint32_t v, uint32_t last) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i + 1 == n ? last : v;
}

// SID order within runs of one key: after the position order (stable: ties
// in creation order), the thread at the start of each run sorts the run by
// (kind, length, alternative sequence, creation order) and flags the first
// record of each SID (runs are short: several records share a position only
// without the infinite-sites model or among drivers and germline mutations).
__device__ __forceinline__ uint32_t sid_rest(const psx_mutation& r) {
    const uint32_t payload = r.kind == PSX_SNV ? r.alt : (r.kind == PSX_INS ? r.seq : 0u);
    return ((uint32_t)r.kind << 24) | ((uint32_t)(r.len & 15u) << 20) | (payload & 0xFFFFFu);
}

__global__ void __launch_bounds__(256) k_sid_runs(const psx_mutation* __restrict__ recs,
                                                  const uint64_t* __restrict__ key, uint32_t* __restrict__ ids,
                                                  uint64_t M, uint8_t* __restrict__ head) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M || (j > 0 && key[j - 1] == key[j])) return;
    uint64_t e = j + 1;
    while (e < M && key[e] == key[j]) ++e;
    if (e - j == 1) {
        head[j] = 1;
        return;
    }
    for (uint64_t i = j + 1; i < e; ++i) {  // insertion sort by (rest, id)
        const uint32_t id = ids[i], ri = sid_rest(recs[id]);
        uint64_t k = i;
        while (k > j) {
            const uint32_t idp = ids[k - 1], rp = sid_rest(recs[idp]);
            if (rp < ri || (rp == ri && idp < id)) break;
            ids[k] = idp;
            --k;
        }
        ids[k] = id;
    }
    uint32_t prev = 0;
    for (uint64_t i = j; i < e; ++i) {
        const uint32_t r = sid_rest(recs[ids[i]]);
        head[i] = (i == j || r != prev) ? 1 : 0;
        prev = r;
    }
}

// Every job's records in SID order into its table, and the kernels' arrays
// of them (the info word as the sequencing kernels read it: allele | ref << 8
// | alt << 10 | kind << 12 | nature << 14 | SID head << 16 | alone at its
// position << 17).
__global__ void __launch_bounds__(256) k_gather(const batch B, const uint64_t* __restrict__ key,
                                                const uint32_t* __restrict__ gids, const uint8_t* __restrict__ head,
                                                uint64_t n) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    psx_mutation r = B.recs[gids[t]];
    r.sid_head = head[t];
    const uint64_t kt = key[t];
    per_job(job_of_seq(B.mbase, B.nj, t), [&](uint32_t j) {
        const uint64_t lo = B.mbase[j], hi = B.mbase[j + 1], i = t - lo;
        const fjob& J = cjob(B, j);
        J.out[i] = r;
        J.pos[i] = r.pos;
        J.lo[i] = r.leaf_lo;
        J.hi[i] = r.leaf_hi;
        const bool solo = (t == lo || key[t - 1] != kt) && (t + 1 == hi || key[t + 1] != kt);
        J.info[i] = uint32_t(r.allele) | (uint32_t(r.ref & 3u) << 8) | (uint32_t(r.alt & 3u) << 10) |
                    (uint32_t(r.kind & 3u) << 12) | (uint32_t(r.nature & 3u) << 14) |
                    (uint32_t(r.sid_head & 1u) << 16) | (uint32_t(solo) << 17);
    });
}

struct index_args {
    uint32_t n_chr;
    uint64_t nb;                              // entries per table
    uint64_t boff[PSX_MAX_CHROMOSOMES + 1];  // first entry of each chromosome
};

// Every table's stratum index: bucket[boff[c] + k] = first record of
// chromosome c at or after position k * PSX_STRATUM (a record index).
__global__ void __launch_bounds__(256) k_index(const batch B, const index_args a) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.nb * B.nj) return;
    const uint32_t j = (uint32_t)(g / a.nb);
    const uint64_t q = g - (uint64_t)j * a.nb;
    const fjob& J = B.jobs[j];
    uint32_t c = 0;
    while (q >= a.boff[c + 1]) ++c;
    const uint64_t target = (q - a.boff[c]) * PSX_STRATUM;
    uint64_t lo = J.chr_begin[c], hi = J.chr_begin[c + 1];
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((uint64_t)J.pos[mid] < target) lo = mid + 1; else hi = mid;
    }
    J.bucket[q] = (uint32_t)lo;
}

// Chromosome ranges of every job's table: first record of chromosome c in its
// position order (c < n_chr), its size beyond; also into a readback array.
__global__ void __launch_bounds__(256) k_chr_begin(const batch B, const uint64_t* __restrict__ key,
                                                   uint64_t* __restrict__ readback) {
    constexpr uint32_t W = PSX_MAX_CHROMOSOMES + 1;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B.nj * W) return;
    const uint32_t j = t / W, c = t - j * W;
    const uint64_t lo = B.mbase[j], M = B.mbase[j + 1] - lo;
    const fjob& J = B.jobs[j];
    const uint64_t v = c < J.n_chr ? lower_bound_u64(key + lo, M, ((uint64_t)j << B.pbits) + B.pstart[c]) : M;
    J.chr_begin[c] = v;
    readback[t] = v;
}

// Bits that represent x (at least 1).
inline int bits_for(uint64_t x) {
    int b = 1;
    while (b < 64 && (x >> b) != 0) ++b;
    return b;
}

inline uint32_t grid_of(uint64_t n) { return (uint32_t)((n + 255) / 256); }
inline uint32_t grid_append(uint64_t n) { return (uint32_t)((n + 256 * APPEND_ITEMS - 1) / (256 * APPEND_ITEMS)); }

inline uint64_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Slots of the moved mutations' table at its largest: at most half full with
// every mutation moved and as many entries left behind (then cleared and
// refilled), and room for the sorted moved positions (12 bytes each) after
// the rounds.
inline uint64_t tab_max(uint64_t M) { return next_pow2(std::max<uint64_t>(1024, 4 * M)); }

// Byte offsets of 256-aligned parts of one buffer.
struct carver {
    size_t bytes = 0;
    size_t take(size_t n) {
        const size_t off = bytes;
        bytes += (std::max<size_t>(n, 1) + 255) & ~size_t(255);
        return off;
    }
};

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

namespace {

// One forest's placement in a batch: host preparation, then the device phases
// that every job of the batch goes through together, one synchronisation per
// phase (node counts; placement and the first infinite-sites check; each
// later round; SID order).
struct place_job {
    uint32_t k = 0;  // index in the batch
    const psx_forest* f = nullptr;
    const psx_engine_params* p = nullptr;
    uint64_t seed = 0;
    psx_mutations* mt = nullptr;
    // host preparation
    std::vector<uint64_t> T;
    size_t expo_off = 0, cls_off = 0, len_off = 0, id_expo_off = 0, id_cls_off = 0;
    std::vector<uint32_t> sk0, soff, sn, ik0, ioff, in_, epoch, drv_d, drv_v;
    std::vector<uint64_t> drv_pos;
    bool use_id = false;
    const psx_id_index* idx = nullptr;
    const psx_sbs_index* sidx = nullptr;
    std::vector<uint64_t> check_pos;  // global positions of the drivers, then the germline mutations
    // scratch offsets
    size_t o_T = 0, o_sk0 = 0, o_soff = 0, o_sn = 0, o_ik0 = 0, o_ioff = 0, o_in = 0, o_epoch = 0, o_drvd = 0,
           o_drvv = 0, o_drivers = 0, o_germ = 0, o_idc = 0, o_idl = 0, o_ido = 0, o_drvpos = 0;
    // sizes and bases in the batch's global arrays
    uint64_t n_pass = 0, n_fixed = 0, M = 0, mbase = 0, pbase = 0;
    place_args pa{};
};

[[noreturn]] void job_throw(const place_job& J, uint32_t n_jobs, psx_status st, const std::string& msg) {
    if (n_jobs > 1) psx_throw(st, "job %u: %s", J.k, msg.c_str());
    psx_throw(st, "%s", msg.c_str());
}

// Validation and the host-side tables of one job (the copy-number events go
// straight into its table); the genome's context and indel indexes are built
// here on first use.
void prepare_job(psx_ctx* ctx, const psx_genome* g, place_job& J) {
    const psx_engine_params* p = J.p;
    const psx_forest* f = J.f;
    const uint64_t seed = J.seed;
    if (p->ploidy < 1 || p->ploidy > 255) psx_throw(PSX_EINVAL, "invalid ploidy");
    if (p->indel_max_len < 1 || p->indel_max_len > 10) psx_throw(PSX_EINVAL, "indel_max_len must be in [1, 10]");
    if (!(p->indel_len_p > 0.0 && p->indel_len_p <= 1.0)) psx_throw(PSX_EINVAL, "invalid indel length parameter");
    if (f->n_nodes && f->max_species >= p->n_species)
        psx_throw(PSX_EINVAL, "node species %u has no rates", f->max_species);
    bool need_snv = p->n_preneoplastic_snv > 0;
    for (uint32_t sp = 0; sp < p->n_species; ++sp) need_snv |= p->snv_rate[sp] > 0.0;
    if (need_snv && (p->n_sbs == 0 || p->n_epochs == 0)) psx_throw(PSX_EINVAL, "SNVs need a signature and an exposure");
    if (p->n_preneoplastic_snv > 0 && p->preneoplastic_sbs >= p->n_sbs)
        psx_throw(PSX_EINVAL, "unknown preneoplastic signature");
    // indels from ID signatures: the genome's indel index (built once per genome) and the class tables
    if (p->n_id) {
        bool need_id = p->n_preneoplastic_indel > 0;
        for (uint32_t sp = 0; sp < p->n_species; ++sp) need_id |= p->indel_rate[sp] > 0.0;
        if (!p->id || (need_id && (!p->id_exposure || p->n_epochs == 0)))
            psx_throw(PSX_EINVAL, "indels need an ID signature and an exposure");
        if (p->n_preneoplastic_indel > 0 && p->preneoplastic_id >= p->n_id)
            psx_throw(PSX_EINVAL, "unknown preneoplastic ID signature");
        if (need_id) {
            J.use_id = true;
            J.idx = &psx_genome_id_index(ctx, const_cast<psx_genome*>(g));
        }
    }
    const uint64_t G = g->boff[g->n_chr];
    // ---- integer tables (host, from the double pmfs) -> one u64 array
    std::vector<uint64_t>& T = J.T;
    auto append = [&](const uint64_t* src, size_t n) {
        const size_t off = T.size();
        T.insert(T.end(), src, src + n);
        return off;
    };
    std::vector<uint64_t> tmp;
    J.expo_off = T.size();
    for (uint32_t e = 0; e < p->n_epochs; ++e) {
        tmp.assign(p->n_sbs, 0);
        psx_cdf_table(p->exposure + (size_t)e * p->n_sbs, p->n_sbs, tmp.data());
        append(tmp.data(), tmp.size());
    }
    // SBS classes: the genome's context index at the engine's sampling (built once per genome and sampling),
    // every class whose context has no indexed site given probability 0
    J.sidx = need_snv ? &psx_genome_sbs_index(ctx, const_cast<psx_genome*>(g), p->context_sampling) : nullptr;
    J.cls_off = T.size();
    for (uint32_t sg = 0; sg < p->n_sbs && J.sidx; ++sg) {
        double pmf[96];
        double mass = 0.0;
        for (uint32_t k = 0; k < 96; ++k) {
            pmf[k] = J.sidx->kept[(k >= 48 ? 16u : 0u) | (k & 15u)] ? p->sbs[(size_t)sg * 96 + k] : 0.0;
            mass += pmf[k];
        }
        if (!(mass > 0.0)) psx_throw(PSX_EINVAL, "SBS signature %u has no class with a site in the context index", sg);
        tmp.assign(96, 0);
        psx_cdf_table(pmf, 96, tmp.data());
        append(tmp.data(), tmp.size());
    }
    std::vector<double> lp(p->indel_max_len);
    for (uint32_t l = 1; l <= p->indel_max_len; ++l)
        lp[l - 1] = std::pow(1.0 - p->indel_len_p, (double)(l - 1)) * p->indel_len_p;
    tmp.assign(p->indel_max_len, 0);
    psx_cdf_table(lp.data(), p->indel_max_len, tmp.data());
    J.len_off = append(tmp.data(), tmp.size());
    if (J.use_id) {
        J.id_expo_off = T.size();
        for (uint32_t e = 0; e < p->n_epochs; ++e) {
            tmp.assign(p->n_id, 0);
            psx_cdf_table(p->id_exposure + (size_t)e * p->n_id, p->n_id, tmp.data());
            append(tmp.data(), tmp.size());
        }
        J.id_cls_off = T.size();
        for (uint32_t sg = 0; sg < p->n_id; ++sg) {
            double pmf[PSX_ID_CLASSES];
            double mass = 0.0;
            for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) {
                pmf[k] = J.idx->count[k] ? p->id[(size_t)sg * PSX_ID_CLASSES + k] : 0.0;
                mass += pmf[k];
            }
            if (!(mass > 0.0)) psx_throw(PSX_EINVAL, "ID signature %u has no class with a site in the reference", sg);
            tmp.assign(PSX_ID_CLASSES, 0);
            psx_cdf_table(pmf, PSX_ID_CLASSES, tmp.data());
            append(tmp.data(), tmp.size());
        }
    }
    const double allelic = (double)((uint64_t)p->ploidy * G);
    J.sk0.resize(p->n_species); J.soff.resize(p->n_species); J.sn.resize(p->n_species);
    J.ik0.resize(p->n_species); J.ioff.resize(p->n_species); J.in_.resize(p->n_species);
    for (uint32_t sp = 0; sp < p->n_species; ++sp) {
        psx_itable a = psx_poisson_table(p->snv_rate[sp] * allelic);
        J.sk0[sp] = a.k0; J.sn[sp] = (uint32_t)a.T.size(); J.soff[sp] = (uint32_t)append(a.T.data(), a.T.size());
        psx_itable b = psx_poisson_table(p->indel_rate[sp] * allelic);
        J.ik0[sp] = b.k0; J.in_[sp] = (uint32_t)b.T.size(); J.ioff[sp] = (uint32_t)append(b.T.data(), b.T.size());
    }
    // ---- drivers and germline (their reference bases are checked after the first device phase), node epochs
    for (uint32_t d = 0; d < p->n_drivers; ++d) {
        const psx_driver& dr = p->drivers[d];
        if (dr.chr >= g->n_chr || dr.pos >= g->len[dr.chr] || dr.alt > 3 ||
            (dr.allele >= 0 && (uint32_t)dr.allele >= p->ploidy))
            psx_throw(PSX_EINVAL, "invalid driver %u", d);
        J.check_pos.push_back(g->boff[dr.chr] + dr.pos);
        for (uint32_t v = 0; v < f->n_nodes; ++v) {
            const uint32_t par = f->parent[v];
            if (f->species[v] == dr.species && (par == PSX_NO_NODE || f->species[par] != dr.species)) {
                J.drv_d.push_back(d);
                J.drv_v.push_back(v);
            }
        }
    }
    // driver loci: no SNV site is drawn there (src/mutation_engine.cpp:244-254)
    for (uint32_t d = 0; d < p->n_drivers; ++d) J.drv_pos.push_back(g->boff[p->drivers[d].chr] + p->drivers[d].pos);
    std::sort(J.drv_pos.begin(), J.drv_pos.end());
    J.drv_pos.erase(std::unique(J.drv_pos.begin(), J.drv_pos.end()), J.drv_pos.end());
    for (uint32_t k = 0; k < p->n_germline; ++k) {
        const psx_germline& gm = p->germline[k];
        if (gm.chr >= g->n_chr || gm.pos >= g->len[gm.chr] || gm.alt > 3 || gm.allele >= p->ploidy)
            psx_throw(PSX_EINVAL, "invalid germline mutation %u", k);
        J.check_pos.push_back(g->boff[gm.chr] + gm.pos);
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
        if (p->n_cna_catalogue == 0 || !p->cna_catalogue) psx_throw(PSX_EINVAL, "passenger CNAs need a CNA catalogue");
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
    J.mt->cna = std::move(cna);
    // epoch of each node's birth (none needed with a single epoch: every node is in epoch 0)
    if (p->n_epochs > 1) {
        J.epoch.assign(f->n_nodes, 0);
        for (uint32_t v = 0; v < f->n_nodes; ++v)
            for (uint32_t e = 0; e < p->n_epochs; ++e)
                if (p->epoch_start[e] <= f->birth[v]) J.epoch[v] = e;
    }
}

void put_job(scratch_builder& sb, place_job& J) {
    const psx_engine_params* p = J.p;
    J.o_T = sb.put_vec(J.T);
    J.o_sk0 = sb.put_vec(J.sk0); J.o_soff = sb.put_vec(J.soff); J.o_sn = sb.put_vec(J.sn);
    J.o_ik0 = sb.put_vec(J.ik0); J.o_ioff = sb.put_vec(J.ioff); J.o_in = sb.put_vec(J.in_);
    J.o_epoch = sb.put_vec(J.epoch);
    J.o_drvd = sb.put_vec(J.drv_d); J.o_drvv = sb.put_vec(J.drv_v);
    J.o_drivers = sb.put(p->drivers, sizeof(psx_driver) * p->n_drivers);
    J.o_germ = sb.put(p->germline, sizeof(psx_germline) * p->n_germline);
    if (J.use_id) {
        J.o_idc = sb.put(J.idx->count, sizeof J.idx->count);
        J.o_idl = sb.put(J.idx->listed, sizeof J.idx->listed);
        J.o_ido = sb.put(J.idx->off, sizeof J.idx->off);
    }
    J.o_drvpos = sb.put_vec(J.drv_pos);
}

// reference bases of a list of global positions (driver and germline checks)
__global__ void k_bases_at(const uint32_t* __restrict__ packed, const uint64_t* __restrict__ woff,
                           const uint64_t* __restrict__ boff, uint32_t n_chr, const uint64_t* __restrict__ gpos,
                           uint32_t n, uint8_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t x = gpos[i];
    uint32_t c = 0;
    while (c + 1 < n_chr && boff[c + 1] <= x) ++c;
    const uint64_t q = x - boff[c];
    out[i] = (uint8_t)((packed[woff[c] + (q >> 4)] >> (2u * (uint32_t)(q & 15u))) & 3u);
}

void place_batch(psx_ctx* ctx, const psx_genome* g, std::vector<place_job>& jobs) {
    hipStream_t s = ctx->stream;
    const uint32_t nj = (uint32_t)jobs.size();
    for (place_job& J : jobs) {
        try {
            prepare_job(ctx, g, J);
        } catch (const psx_error& e) {
            job_throw(J, nj, e.code, e.msg);
        }
    }
    const uint64_t G = g->boff[g->n_chr];
    // ---- descriptors of every job in one scratch upload
    scratch_builder sb;
    const size_t o_len = sb.put_vec(g->len), o_woff = sb.put_vec(g->woff), o_boff = sb.put_vec(g->boff);
    std::vector<uint64_t> pstart(g->n_chr);  // padded chromosome starts: boff[c] + 16 c
    for (uint32_t c = 0; c < g->n_chr; ++c) pstart[c] = g->boff[c] + 16ull * c;
    const size_t o_pstart = sb.put_vec(pstart);
    std::vector<uint64_t> check_pos;
    for (place_job& J : jobs) {
        put_job(sb, J);
        check_pos.insert(check_pos.end(), J.check_pos.begin(), J.check_pos.end());
    }
    const size_t o_check = sb.put_vec(check_pos);
    uint8_t* base = (uint8_t*)psx_scratch_upload(ctx, sb);
    auto at = [&](size_t off) { return (void*)(base + off); };
    // sort keys: the padded position below bit pbits (at least 16 to spare), the job above
    const uint64_t p_max = pstart[g->n_chr - 1] + g->len[g->n_chr - 1];
    const uint32_t pbits = (uint32_t)bits_for(p_max + 16), kbits = pbits + (uint32_t)bits_for(nj - 1);
    if (kbits > 64) psx_throw(PSX_EINVAL, "too many jobs for one call (%u)", nj);

    // node bases, the job table and the batch's round state (one device buffer)
    std::vector<uint64_t> nbase(nj + 1, 0), mbase(nj + 1, 0);
    for (uint32_t k = 0; k < nj; ++k) nbase[k + 1] = nbase[k] + jobs[k].f->n_nodes;
    const uint64_t N = nbase[nj];
    carver ct;
    const size_t t_jobs = ct.take(sizeof(fjob) * nj), t_nbase = ct.take(8 * (nj + 1)),
                 t_mbase = ct.take(8 * (nj + 1)), t_state = ct.take(sizeof(round_io) + 4 * nj);
    ctx->place_tab.reserve(ct.bytes);
    uint8_t* dt = (uint8_t*)ctx->place_tab.p;
    round_io* d_io = (round_io*)(dt + t_state);
    uint32_t* d_status = (uint32_t*)(d_io + 1);
    batch B{};
    B.jobs = (const fjob*)(dt + t_jobs);
    B.nj = nj;
    B.pbits = pbits;
    B.nbase = (const uint64_t*)(dt + t_nbase);
    B.mbase = (const uint64_t*)(dt + t_mbase);
    B.pstart = (const uint64_t*)at(o_pstart);

    // pinned host area: the readbacks (the round state and every job's status, the jobs' passenger counts,
    // their chromosome ranges, the checked reference bases), then the job table staged in the device layout
    constexpr uint32_t W = PSX_MAX_CHROMOSOMES + 1;
    const size_t io_state = sizeof(round_io) + 4 * nj;
    const size_t r_pcount = (io_state + 7) / 8, r_chr = r_pcount + nj, r_bases = r_chr + (size_t)nj * W;
    const size_t r_stage = r_bases + (check_pos.size() + 8 + 255) / 8 / 32 * 32 + 32;  // past the bases
    uint64_t* io = (uint64_t*)psx_host_io(ctx, 8 * r_stage + t_state);
    const round_io* h_io = (const round_io*)io;
    const uint32_t* h_status = (const uint32_t*)(h_io + 1);
    uint8_t* io_bases = (uint8_t*)(io + r_bases);
    uint8_t* stage = (uint8_t*)(io + r_stage);
    fjob* tab = (fjob*)(stage + t_jobs);
    for (uint32_t k = 0; k < nj; ++k) tab[k] = fjob{};
    // the job table, bases included, in one copy (the stream is idle or in order before each use)
    auto upload_table = [&] {
        std::memcpy(stage + t_nbase, nbase.data(), 8 * (nj + 1));
        std::memcpy(stage + t_mbase, mbase.data(), 8 * (nj + 1));
        PSX_HIP(hipMemcpyAsync(dt, stage, t_state, hipMemcpyHostToDevice, s));
    };

    PSX_HIP(hipEventRecord(ctx->ev0, s));
    // ---- phase 1: passenger counts per node of every job, the driver and germline reference bases
    size_t scan_bytes = 0;
    PSX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (int)(N + 1), s));
    carver cn;
    const size_t c_scan = cn.take(scan_bytes), c_bases = cn.take(check_pos.size()), c_cs = cn.take(N * 4),
                 c_ct = cn.take((N + 1) * 8), c_off = cn.take((N + 1) * 8), c_offj = cn.take(N * 8),
                 c_pc = cn.take(nj * 8), c_cb = cn.take((size_t)nj * W * 8);
    ctx->place_nodes.reserve(cn.bytes);
    uint8_t* pn = (uint8_t*)ctx->place_nodes.p;
    uint32_t* d_cnt_snv = (uint32_t*)(pn + c_cs);
    uint64_t *d_cnt_tot = (uint64_t*)(pn + c_ct), *d_off = (uint64_t*)(pn + c_off),
             *d_off_job = (uint64_t*)(pn + c_offj), *d_pcount = (uint64_t*)(pn + c_pc),
             *d_cb = (uint64_t*)(pn + c_cb);
    if (!check_pos.empty()) {
        hipLaunchKernelGGL(k_bases_at, dim3(grid_of(check_pos.size())), dim3(256), 0, s, g->packed.as<uint32_t>(),
                           (const uint64_t*)at(o_woff), (const uint64_t*)at(o_boff), g->n_chr,
                           (const uint64_t*)at(o_check), (uint32_t)check_pos.size(), pn + c_bases);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipMemcpyAsync(io_bases, pn + c_bases, check_pos.size(), hipMemcpyDeviceToHost, s));
    }
    for (uint32_t k = 0; k < nj; ++k) {
        place_job& J = jobs[k];
        edge_args& ea = tab[k].ea;
        ea.seed = J.seed; ea.n_nodes = J.f->n_nodes; ea.parent = J.f->d_parent.as<uint32_t>();
        ea.species = J.f->d_species.as<uint32_t>();
        ea.snv_k0 = (uint32_t*)at(J.o_sk0); ea.snv_off = (uint32_t*)at(J.o_soff); ea.snv_n = (uint32_t*)at(J.o_sn);
        ea.ind_k0 = (uint32_t*)at(J.o_ik0); ea.ind_off = (uint32_t*)at(J.o_ioff); ea.ind_n = (uint32_t*)at(J.o_in);
        ea.T = (uint64_t*)at(J.o_T); ea.cnt_snv = d_cnt_snv + nbase[k]; ea.cnt_tot = d_cnt_tot + nbase[k];
    }
    upload_table();
    PSX_HIP(hipMemsetAsync(d_cnt_tot + N, 0, 8, s));
    if (N) {
        hipLaunchKernelGGL(k_edge_counts, dim3(grid_of(N)), dim3(256), 0, s, B, N);
        PSX_HIP(hipGetLastError());
    }
    PSX_HIP(hipcub::DeviceScan::ExclusiveSum(pn + c_scan, scan_bytes, d_cnt_tot, d_off, (int)(N + 1), s));
    hipLaunchKernelGGL(k_node_bases, dim3(grid_of(std::max<uint64_t>(N, nj))), dim3(256), 0, s, B, N, d_off,
                       d_off_job, d_pcount);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipMemcpyAsync(io + r_pcount, d_pcount, 8 * nj, hipMemcpyDeviceToHost, s));
    PSX_HIP(hipStreamSynchronize(s));
    {
        size_t ci = 0;
        for (place_job& J : jobs) {
            for (uint32_t d = 0; d < J.p->n_drivers; ++d, ++ci)
                if (J.p->drivers[d].alt == io_bases[ci]) job_throw(J, nj, PSX_EINVAL, "invalid driver " + std::to_string(d));
            for (uint32_t q = 0; q < J.p->n_germline; ++q, ++ci)
                if (J.p->germline[q].alt == io_bases[ci])
                    job_throw(J, nj, PSX_EINVAL, "invalid germline mutation " + std::to_string(q));
        }
    }
    // ---- phase 2: mutation ids, placement and the first infinite-sites check of every job, one sort
    uint64_t n_pass_all = 0;
    for (uint32_t k = 0; k < nj; ++k) {
        place_job& J = jobs[k];
        J.n_pass = io[r_pcount + k];
        J.n_fixed = J.drv_d.size() + J.p->n_germline;
        J.M = J.n_fixed + J.p->n_preneoplastic_snv + J.p->n_preneoplastic_indel + J.n_pass;
        if (J.M > ID_MASK)
            job_throw(J, nj, PSX_EINVAL, "too many mutations (" + std::to_string(J.M) + ", at most " +
                                             std::to_string(ID_MASK) + ")");
        J.mbase = mbase[k];
        J.pbase = n_pass_all;
        mbase[k + 1] = mbase[k] + J.M;
        n_pass_all += J.n_pass;
    }
    const uint64_t Mt = mbase[nj];
    if (Mt >= NO_ENTRY) psx_throw(PSX_EINVAL, "too many mutations in one call (%llu)", (unsigned long long)Mt);
    size_t sort_bytes = 0, rank_bytes = 0;
    PSX_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                               (uint32_t*)nullptr, (uint32_t*)nullptr, (int)std::max<uint64_t>(Mt, 1),
                                               0, (int)kbits, s));
    PSX_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, rank_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (int)Mt + 1, s));
    sort_bytes = std::max(sort_bytes, rank_bytes);
    // per-mutation arrays of the whole batch from the context's placement pool: attempt counters, moved flags
    // and loser stamps first (cleared together)
    const uint64_t tab_slots_max = tab_max(Mt);
    carver cm;
    const size_t c_zero = cm.take(Mt * 12), c_sort = cm.take(sort_bytes), c_recs = cm.take(Mt * sizeof(psx_mutation)),
                 c_list0 = cm.take(Mt * 4), c_list1 = cm.take(Mt * 4), c_lap0 = cm.take(Mt * 4),
                 c_lap1 = cm.take(Mt * 4), c_ids = cm.take(Mt * 4), c_ids2 = cm.take(Mt * 4), c_ml = cm.take(Mt * 4),
                 c_map = cm.take(Mt * 4), c_w = cm.take((Mt + 1) * 4), c_S = cm.take((Mt + 1) * 4),
                 c_nof = cm.take(n_pass_all * 4), c_key = cm.take(Mt * 8), c_key2 = cm.take(Mt * 8),
                 c_tab = cm.take(tab_slots_max * 4), c_head = cm.take(Mt);
    ctx->place_muts.reserve(cm.bytes);
    uint8_t* pm = (uint8_t*)ctx->place_muts.p;
    void* d_sort = pm + c_sort;
    uint32_t* d_att = (uint32_t*)(pm + c_zero);
    uint32_t* d_moved = d_att + Mt;
    uint32_t* d_lose = d_att + 2 * Mt;
    psx_mutation* d_recs = (psx_mutation*)(pm + c_recs);
    uint32_t* d_list[2] = {(uint32_t*)(pm + c_list0), (uint32_t*)(pm + c_list1)};
    uint32_t* d_lapos[2] = {(uint32_t*)(pm + c_lap0), (uint32_t*)(pm + c_lap1)};
    uint32_t *d_ids = (uint32_t*)(pm + c_ids), *d_ids2 = (uint32_t*)(pm + c_ids2), *d_mlist = (uint32_t*)(pm + c_ml),
             *d_mapos = (uint32_t*)(pm + c_map), *d_w = (uint32_t*)(pm + c_w), *d_S = (uint32_t*)(pm + c_S),
             *d_node_of = (uint32_t*)(pm + c_nof), *d_tab = (uint32_t*)(pm + c_tab);
    uint64_t *d_key = (uint64_t*)(pm + c_key), *d_key2 = (uint64_t*)(pm + c_key2);
    uint8_t* d_head = pm + c_head;
    B.recs = d_recs;
    for (uint32_t k = 0; k < nj; ++k) {
        place_job& J = jobs[k];
        const psx_engine_params* p = J.p;
        place_args& pa = tab[k].pa;
        pa.g = gdev{g->packed.as<uint32_t>(), (uint32_t*)at(o_len), (uint64_t*)at(o_woff), (uint64_t*)at(o_boff), G,
                    g->n_chr};
        pa.seed = J.seed; pa.ploidy = p->ploidy; pa.n_sbs = p->n_sbs;
        pa.expo_T = (uint64_t*)at(J.o_T) + J.expo_off; pa.cls_T = (uint64_t*)at(J.o_T) + J.cls_off;
        pa.len_T = (uint64_t*)at(J.o_T) + J.len_off; pa.len_n = p->indel_max_len;
        pa.node_epoch = J.epoch.empty() ? nullptr : (uint32_t*)at(J.o_epoch);
        pa.leaf_lo = J.f->d_leaf_lo.as<uint32_t>(); pa.leaf_hi = J.f->d_leaf_hi.as<uint32_t>();
        pa.n_leaves = J.f->n_leaves; pa.n_nodes = J.f->n_nodes;
        pa.n_drv = J.drv_d.size(); pa.n_germ = p->n_germline;
        pa.n_pre_snv = p->n_preneoplastic_snv; pa.n_pre_indel = p->n_preneoplastic_indel;
        pa.drv_d = (uint32_t*)at(J.o_drvd); pa.drv_v = (uint32_t*)at(J.o_drvv);
        pa.drivers = (psx_driver*)at(J.o_drivers); pa.germ = (psx_germline*)at(J.o_germ);
        pa.pre_sig = p->preneoplastic_sbs;
        pa.cnt_snv = d_cnt_snv + nbase[k]; pa.node_off = d_off_job + nbase[k]; pa.node_of = d_node_of + J.pbase;
        pa.M = J.M; pa.recs = d_recs + J.mbase; pa.attempt = d_att + J.mbase; pa.status = d_status + k;
        if (J.sidx) {
            pa.sbs_list = J.sidx->d_list.as<uint32_t>();
            pa.sbs_strand = J.sidx->strand_bit ? 1u : 0u;
            pa.sbs_kept = J.sidx->d_meta.as<uint64_t>();
            pa.sbs_off = J.sidx->d_meta.as<uint64_t>() + 32;
        }
        pa.drv_pos = (const uint64_t*)at(J.o_drvpos);
        pa.n_drv_pos = (uint32_t)J.drv_pos.size();
        if (J.use_id) {
            pa.use_id = 1; pa.n_id = p->n_id; pa.pre_id = p->preneoplastic_id;
            pa.id_expo_T = (uint64_t*)at(J.o_T) + J.id_expo_off; pa.id_cls_T = (uint64_t*)at(J.o_T) + J.id_cls_off;
            pa.id_count = (const uint64_t*)at(J.o_idc); pa.id_listed = (const uint32_t*)at(J.o_idl);
            pa.id_off = (const uint64_t*)at(J.o_ido); pa.id_list = J.idx->d_list.as<uint64_t>();
        }
        tab[k].mbase = J.mbase;
        tab[k].n_fixed = J.n_fixed;
        tab[k].is = p->infinite_sites ? 1u : 0u;
        tab[k].n_chr = g->n_chr;
    }
    upload_table();
    PSX_HIP(hipMemsetAsync(dt + t_state, 0, sizeof(round_io) + 4 * nj, s));
    PSX_HIP(hipMemsetAsync(pm + c_zero, 0, Mt * 12, s));
    if (N) {
        hipLaunchKernelGGL(k_node_expand, dim3(grid_of(N)), dim3(256), 0, s, B, N, d_off, d_node_of);
        PSX_HIP(hipGetLastError());
    }
    if (Mt) {
        // first attempts and sort entries; one sort of every job's footprints (job-major, the sorted entries stay
        // as the reference of later rounds and give the position order); the first check
        hipLaunchKernelGGL(k_place, dim3(grid_of(Mt)), dim3(256), 0, s, B, Mt, d_key, d_ids);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipcub::DeviceRadixSort::SortPairs(d_sort, sort_bytes, d_key, d_key2, d_ids, d_ids2, (int)Mt, 0,
                                                   (int)kbits, s));
        hipLaunchKernelGGL(k_conflicts, dim3(grid_append(Mt)), dim3(256), 0, s, B, d_key2, d_ids2, Mt, d_lose,
                           d_list[0], d_lapos[0], &d_io->n_lose[0]);
        PSX_HIP(hipGetLastError());
    }
    PSX_HIP(hipMemcpyAsync(io, dt + t_state, io_state, hipMemcpyDeviceToHost, s));
    // the tables' block, allocated while the GPU runs the placement
    size_t block_bytes = 0;
    std::vector<size_t> t_off(nj);
    for (uint32_t k = 0; k < nj; ++k) {
        t_off[k] = block_bytes;
        block_bytes += psx_mutations_bytes(jobs[k].M);
    }
    auto block = std::make_shared<dbuf>();
    block->alloc(block_bytes);
    PSX_HIP(hipStreamSynchronize(s));
    // status of every job's placement launches, the moved mutations and the losers of the last check run
    uint32_t round = 1, spec_end = 1, n_lose = 0, n_moved = 0;
    bool halted = false;
    auto take_status = [&] {
        for (place_job& J : jobs) {
            const uint32_t st = h_status[J.k];
            if (st & ST_NO_SITE) job_throw(J, nj, PSX_EINVAL, "no reference site matches a drawn SBS or ID class");
            if (st & ST_NO_INDEL_ROOM) job_throw(J, nj, PSX_EINVAL, "no room for an indel");
        }
        n_moved = h_io->n_moved;
        if (h_io->halt) {  // resumed at the halted round once the table is rebuilt
            round = h_io->halt;
            halted = true;
        } else {
            round = spec_end;
            halted = false;
        }
        n_lose = h_io->n_lose[round - 1];
    };
    take_status();
    // ---- phase 3: later infinite-sites rounds of every job together, SPEC_ROUNDS of them launched between
    // synchronisations (rounds after the last loser do nothing)
    uint64_t tab_slots = 0;
    while (n_lose) {
        if (round >= (uint32_t)IS_ROUNDS) {
            // the job of a remaining loser
            uint32_t gid = 0;
            PSX_HIP(hipMemcpy(&gid, d_list[(round - 1) & 1], 4, hipMemcpyDeviceToHost));
            uint32_t k = 0;
            while (k + 1 < nj && mbase[k + 1] <= gid) ++k;
            job_throw(jobs[k], nj, PSX_EINVAL, "infinite-sites placement did not converge");
        }
        if (!tab_slots || halted) {
            // cleared, sized for a few rounds of this many losers, refilled with the moved footprints
            tab_slots = std::min(tab_slots_max, next_pow2(std::max<uint64_t>(1024, 8 * ((uint64_t)n_moved + n_lose))));
            PSX_HIP(hipMemsetAsync(d_tab, 0xFF, tab_slots * 4, s));
            hipLaunchKernelGGL(k_tab_fill, dim3(std::max<uint32_t>(1, grid_of(n_moved))), dim3(256), 0, s, B, d_mlist,
                               n_moved, d_tab, (uint32_t)(tab_slots - 1), d_io);
            PSX_HIP(hipGetLastError());
        }
        const uint32_t grid = std::min<uint32_t>(1024, grid_of(n_lose));
        spec_end = std::min<uint32_t>(round + SPEC_ROUNDS, IS_ROUNDS);
        for (uint32_t r = round; r < spec_end; ++r) {
            redraw_args rd{};
            rd.round = r; rd.list = d_list[(r - 1) & 1]; rd.lapos = d_lapos[(r - 1) & 1];
            rd.moved = d_moved; rd.mlist = d_mlist; rd.mapos = d_mapos;
            rd.akey = d_key2; rd.aval = d_ids2; rd.n = Mt; rd.io = d_io;
            rd.tab = d_tab; rd.tab_mask = (uint32_t)(tab_slots - 1);
            hipLaunchKernelGGL(k_redraw, dim3(grid), dim3(256), 0, s, B, rd);
            PSX_HIP(hipGetLastError());
            check_args ca{};
            ca.akey = d_key2; ca.aval = d_ids2; ca.n = Mt; ca.tab = d_tab; ca.tab_mask = rd.tab_mask;
            ca.round = r; ca.list = d_list[(r - 1) & 1]; ca.io = d_io;
            ca.lose = d_lose; ca.out = d_list[r & 1]; ca.out_apos = d_lapos[r & 1];
            hipLaunchKernelGGL(k_check, dim3(grid), dim3(256), 0, s, B, ca);
            PSX_HIP(hipGetLastError());
        }
        PSX_HIP(hipMemcpyAsync(io, dt + t_state, io_state, hipMemcpyDeviceToHost, s));
        PSX_HIP(hipStreamSynchronize(s));
        take_status();
    }
    // ---- phase 4: SID order into the tables, which share one block: the first check's sorted entries are in
    // position order but for the moved mutations, which are sorted alone and merged in; then each run of one
    // position is sorted by the rest of the SID
    for (uint32_t k = 0; k < nj; ++k) {
        place_job& J = jobs[k];
        psx_mutations* mt = J.mt;
        mt->n = J.M;
        mt->ploidy = J.p->ploidy;
        mt->n_leaves = J.f->n_leaves;
        mt->chr_len = g->len;
        psx_mutations_carve(mt, block, t_off[k]);
        tab[k].out = mt->recs;
        tab[k].pos = mt->pos; tab[k].lo = mt->lo; tab[k].hi = mt->hi; tab[k].info = mt->info;
        tab[k].chr_begin = mt->d_chr_begin;
    }
    // the tables' stratum indexes for the genome's chromosome lengths, in one block
    const std::vector<uint64_t> ix_off = psx_index_layout(g->len);
    index_args ia{};
    ia.n_chr = g->n_chr;
    ia.nb = ix_off[g->n_chr];
    for (uint32_t c = 0; c <= g->n_chr; ++c) ia.boff[c] = ix_off[c];
    auto ix_block = std::make_shared<dbuf>();
    ix_block->alloc(ia.nb * nj * 4u);
    for (uint32_t k = 0; k < nj; ++k) tab[k].bucket = ix_block->as<uint32_t>() + ia.nb * k;
    upload_table();
    if (Mt) {
        merge_args ma{};
        ma.akey = d_key2; ma.aval = d_ids2; ma.n = Mt; ma.nb = n_moved; ma.okey = d_key; ma.ogid = d_ids;
        if (n_moved) {
            // the sorted moved keys go into the table's memory, free after the rounds, their upper bounds into
            // the first loser list's
            uint64_t* bkey = (uint64_t*)d_tab;
            uint32_t* bval = (uint32_t*)(bkey + n_moved);
            hipLaunchKernelGGL(k_moved_keys, dim3(grid_of(n_moved)), dim3(256), 0, s, B, d_mlist, n_moved, d_key,
                               d_ids);
            PSX_HIP(hipGetLastError());
            PSX_HIP(hipcub::DeviceRadixSort::SortPairs(d_sort, sort_bytes, d_key, bkey, d_ids, bval, (int)n_moved, 0,
                                                       (int)kbits, s));
            hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(Mt + 1)), dim3(256), 0, s, d_w, Mt + 1, 1u, 0u);
            PSX_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_merge_marks, dim3(grid_of(n_moved)), dim3(256), 0, s, d_mapos, d_key2, Mt, bkey,
                               n_moved, d_w, d_list[0]);
            PSX_HIP(hipGetLastError());
            PSX_HIP(hipcub::DeviceScan::ExclusiveSum(d_sort, sort_bytes, d_w, d_S, (int)Mt + 1, s));
            ma.S = d_S; ma.bkey = bkey; ma.bval = bval; ma.ub = d_list[0];
        }
        hipLaunchKernelGGL(k_merge, dim3(grid_of(Mt + n_moved)), dim3(256), 0, s, B, ma);
        PSX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sid_runs, dim3(grid_of(Mt)), dim3(256), 0, s, d_recs, d_key, d_ids, Mt, d_head);
        PSX_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_gather, dim3(grid_of(Mt)), dim3(256), 0, s, B, d_key, d_ids, d_head, Mt);
        PSX_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_chr_begin, dim3(grid_of((uint64_t)nj * W)), dim3(256), 0, s, B, d_key, d_cb);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipMemcpyAsync(io + r_chr, d_cb, (size_t)nj * W * 8, hipMemcpyDeviceToHost, s));
    hipLaunchKernelGGL(k_index, dim3(grid_of(ia.nb * nj)), dim3(256), 0, s, B, ia);
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipEventRecord(ctx->ev1, s));
    PSX_HIP(hipStreamSynchronize(s));
    for (uint32_t k = 0; k < nj; ++k) {
        psx_mutations* mt = jobs[k].mt;
        const uint64_t* cb = io + r_chr + (size_t)k * W;
        mt->chr_begin.assign(cb, cb + W);
        auto ix = std::make_unique<psx_mut_index>();
        ix->chr_len = g->len;
        ix->bucket_off = ix_off;
        ix->store = ix_block;
        ix->bucket = tab[k].bucket;
        std::lock_guard<std::mutex> lk(mt->idx_mu);
        mt->index.push_back(std::move(ix));
    }
    float ms = 0.f;
    PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_place_ms = ms;
}

}  // namespace

extern "C" psx_status psx_place_mutations_batch(psx_ctx* ctx, const psx_genome* g, uint32_t n_jobs,
                                                const psx_forest* const* f, const psx_engine_params* const* p,
                                                const uint64_t* seeds, psx_mutations** out) {
    if (!ctx || !g || !f || !p || !seeds || !out || n_jobs == 0) return PSX_EINVAL;
    for (uint32_t k = 0; k < n_jobs; ++k)
        if (!f[k] || !p[k]) return PSX_EINVAL;
    std::vector<place_job> jobs(n_jobs);
    for (uint32_t k = 0; k < n_jobs; ++k) {
        jobs[k].k = k;
        jobs[k].f = f[k];
        jobs[k].p = p[k];
        jobs[k].seed = seeds[k];
        jobs[k].mt = new psx_mutations();
        jobs[k].mt->ctx = ctx;
    }
    const psx_status st = psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        place_batch(ctx, g, jobs);
    });
    if (st != PSX_OK) {
        for (place_job& J : jobs) delete J.mt;
        return st;
    }
    for (uint32_t k = 0; k < n_jobs; ++k) out[k] = jobs[k].mt;
    return PSX_OK;
}

extern "C" psx_status psx_place_mutations(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                          const psx_engine_params* p, uint64_t seed, psx_mutations** out) {
    if (!ctx || !g || !f || !p || !out) return PSX_EINVAL;
    return psx_place_mutations_batch(ctx, g, 1, &f, &p, &seed, out);
}
