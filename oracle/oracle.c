/*
 * CPU oracle: TEST INFRASTRUCTURE ONLY (see oracle.h). Never linked into the
 * product; loaded by tests/, the smoke check and bench.py's CPU baseline.
 *
 * This file is the executable statement of the sampling model. Each function
 * names the reference code whose behaviour it restates.
 */
#include "oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

enum {
    P_REF = 0x01, P_TEMPLATE = 0x10, P_ERROR = 0x11, P_EDGE = 0x20, P_SNV = 0x21,
    P_SITE = 0x22, P_INDEL = 0x23, P_IDCLS = 0x24, P_DRIVER = 0x25, P_IDSITE = 0x26, P_PCNA = 0x27
};

#define SITE_TRIES (1u << 20)
#define LIST_TRIES 4096u  /* draws of a context-index site before the class is given up as driver loci only */
#define INDEL_TRIES (1u << 20)
#define IS_ROUNDS 256  /* checks of the infinite-sites model before placement gives up */
#define TABLE_CAP (1u << 16)
#define TWO32 4294967296.0

static __thread char g_err[512];

static int fail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return PSX_EINVAL;
}

const char* orc_last_error(void) { return g_err; }
void orc_free(void* p) { free(p); }

int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------ Philox4x32-10 */

void orc_philox(const uint32_t ctr[4], uint64_t key, uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox4(uint64_t key, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t o[4]) {
    uint32_t ctr[4] = {a, b, c, d};
    orc_philox(ctr, key, o);
}

static uint32_t mulhi32(uint32_t u, uint32_t n) { return (uint32_t)(((uint64_t)u * n) >> 32); }

static uint64_t mulhi64(uint32_t hi, uint32_t lo, uint64_t n) {
    unsigned __int128 u = ((unsigned __int128)(((uint64_t)hi << 32) | lo)) * n;
    return (uint64_t)(u >> 64);
}

/* ------------------------------------------------------------ inversion tables */

int orc_cdf_table(const double* pmf, uint32_t n, uint64_t* T) {
    if (n == 0) return fail("empty distribution");
    double total = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
        if (!(pmf[k] >= 0.0)) return fail("negative or NaN probability");
        total += pmf[k];
    }
    if (!(total > 0.0) || !isfinite(total)) return fail("distribution has no mass");
    double cum = 0.0;
    for (uint32_t k = 0; k + 1 < n; ++k) {
        cum += pmf[k];
        double x = cum / total;
        T[k] = x >= 1.0 ? (uint64_t)1 << 32 : (uint64_t)(x * TWO32);
    }
    T[n - 1] = (uint64_t)1 << 32;
    return 0;
}

/* Linear scan (the kernels use a binary search; both give the unique k). */
static uint32_t pick(const uint64_t* T, uint32_t n, uint32_t u) {
    uint32_t k = 0;
    while (k + 1 < n && (uint64_t)u >= T[k]) ++k;
    return k;
}

int orc_poisson_table(double lambda, uint32_t* k0, uint32_t* n, uint64_t* T, uint32_t cap) {
    if (!(lambda >= 0.0) || !isfinite(lambda)) return fail("invalid Poisson mean %g", lambda);
    if (lambda == 0.0) {
        *k0 = 0; *n = 1; T[0] = (uint64_t)1 << 32;
        return 0;
    }
    double sd = sqrt(lambda);
    double lo = floor(lambda - 40.0 * sd - 10.0);
    if (lo < 0.0) lo = 0.0;
    double hi = ceil(lambda + 40.0 * sd + 10.0);
    double cnt = hi - lo + 1.0;
    if (cnt > (double)cap) return fail("Poisson mean %g too large", lambda);
    uint32_t m = (uint32_t)cnt;
    double* pmf = (double*)malloc(sizeof(double) * m);
    double ll = log(lambda);
    for (uint32_t i = 0; i < m; ++i) {
        double k = lo + (double)i;
        pmf[i] = exp(-lambda + k * ll - lgamma(k + 1.0));
    }
    int rc = orc_cdf_table(pmf, m, T);
    free(pmf);
    *k0 = (uint32_t)lo; *n = m;
    return rc;
}

static int binomial_range(uint32_t t, double p, double lo, double hi, uint32_t cap, uint32_t* k0,
                          uint32_t* n, uint64_t* T) {
    double cnt = hi - lo + 1.0;
    if (cnt > (double)cap) return fail("binomial table too large");
    uint32_t m = (uint32_t)cnt;
    double* pmf = (double*)malloc(sizeof(double) * m);
    double lp = log(p), lq = log1p(-p), lgt = lgamma((double)t + 1.0);
    for (uint32_t i = 0; i < m; ++i) {
        double k = lo + (double)i;
        pmf[i] = exp(lgt - lgamma(k + 1.0) - lgamma((double)t - k + 1.0) + k * lp + ((double)t - k) * lq);
    }
    int rc = orc_cdf_table(pmf, m, T);
    free(pmf);
    *k0 = (uint32_t)lo; *n = m;
    return rc;
}

int orc_binomial_table(uint32_t t, double p, uint32_t* k0, uint32_t* n, uint64_t* T, uint32_t cap) {
    if (!(p > 0.0) || p > 1.0) return fail("invalid binomial probability %g", p);
    if (p == 1.0) {
        *k0 = t; *n = 1; T[0] = (uint64_t)1 << 32;
        return 0;
    }
    double m = (double)t * p, sd = sqrt((double)t * p * (1.0 - p));
    double lo = floor(m - 40.0 * sd - 10.0), hi = ceil(m + 40.0 * sd + 10.0);
    if (lo < 0.0) lo = 0.0;
    if (hi > (double)t) hi = (double)t;
    return binomial_range(t, p, lo, hi, cap, k0, n, T);
}

/* Whether a template's N = rpt L bases hold more than PSX_ERR_KCAP errors
 * with probability below 2^-40 (Binomial(N, e) tail summed term by term from
 * PSX_ERR_KCAP + 1). */
static int error_rate_ok(uint32_t N, double e) {
    if (N <= PSX_ERR_KCAP) return 1;
    double lp = log(e), lq = log1p(-e), lgn = lgamma((double)N + 1.0), tail = 0.0;
    for (uint32_t k = PSX_ERR_KCAP + 1; k <= N; ++k) {
        double term = exp(lgn - lgamma((double)k + 1.0) - lgamma((double)(N - k) + 1.0) + k * lp +
                          (double)(N - k) * lq);
        tail += term;
        if (tail >= 0x1p-40 || (term < 1e-300 && (double)k > (double)N * e)) break;
    }
    return tail < 0x1p-40;
}

/* Errors of a template's N bases: Binomial(N, e) on [0, min(N, PSX_ERR_KCAP)]. */
static int error_table(uint32_t N, double e, uint32_t* k0, uint32_t* n, uint64_t* T) {
    double hi = N < PSX_ERR_KCAP ? (double)N : (double)PSX_ERR_KCAP;
    return binomial_range(N, e, 0.0, hi, PSX_ERR_KCAP + 1, k0, n, T);
}

/* get_bin_dist, src/seq_simulation.cpp:431-451: q = sd^2/mean, p = 1 - q,
 * error when p < 0 (p == 0 is rejected as well: t would be infinite),
 * t = (uint32)(mean / p). */
static int insert_table(uint32_t mean, uint32_t sd, uint32_t* k0, uint32_t* n, uint64_t* T) {
    double q = (double)(sd * sd) / (double)mean;
    double p = 1.0 - q;
    if (!(p > 0.0))
        return fail("The insert size mean (%u) must be greater than or equal to its variance "
                    "(%u*%u=%u).", mean, sd, sd, sd * sd);
    uint32_t t = (uint32_t)((double)mean / p);
    return orc_binomial_table(t, p, k0, n, T, 4096);
}

/* -------------------------------------------------------------------- genome */

uint64_t orc_genome_words(uint32_t n_chr, const uint32_t* len) {
    uint64_t w = 0;
    for (uint32_t c = 0; c < n_chr; ++c) w += (len[c] + 15ull) / 16;
    return w;
}

/* Synthetic reference (stands in for the FASTA of src/mutation_engine.cpp:75-112):
 * word w of chromosome c is word (w % 4) of Philox(key = seed,
 * counter = (w / 4, c, 0, REF << 24)); bits past the chromosome end are 0. */
int orc_genome_synth(uint32_t n_chr, const uint32_t* len, uint64_t seed, uint32_t* packed) {
    if (n_chr == 0 || n_chr > PSX_MAX_CHROMOSOMES) return fail("invalid number of chromosomes");
    uint64_t off = 0;
    for (uint32_t c = 0; c < n_chr; ++c) {
        uint64_t words = (len[c] + 15ull) / 16;
        for (uint64_t w = 0; w < words; ++w) {
            uint32_t o[4];
            philox4(seed, (uint32_t)(w / 4), c, 0, (uint32_t)P_REF << 24, o);
            packed[off + w] = o[w % 4];
        }
        uint32_t rem = len[c] % 16;
        if (rem) packed[off + words - 1] &= (1u << (2 * rem)) - 1u;
        off += words;
    }
    return 0;
}

typedef struct genome_view {
    uint32_t n_chr;
    const uint32_t* len;
    const uint32_t* packed;
    uint64_t woff[PSX_MAX_CHROMOSOMES];
    uint64_t boff[PSX_MAX_CHROMOSOMES + 1]; /* absolute base offset of each chromosome */
} genome_view;

static void gv_init(genome_view* g, uint32_t n_chr, const uint32_t* len, const uint32_t* packed) {
    g->n_chr = n_chr; g->len = len; g->packed = packed;
    uint64_t w = 0, b = 0;
    for (uint32_t c = 0; c < n_chr; ++c) {
        g->woff[c] = w; g->boff[c] = b;
        w += (len[c] + 15ull) / 16;
        b += len[c];
    }
    g->boff[n_chr] = b;
}

static uint32_t base_at(const genome_view* g, uint32_t c, uint32_t i) {
    return (g->packed[g->woff[c] + i / 16] >> (2 * (i % 16))) & 3u;
}

static void locate(const genome_view* g, uint64_t x, uint32_t* c, uint32_t* pos) {
    uint32_t k = 0;
    while (x >= g->boff[k + 1]) ++k;
    *c = k;
    *pos = (uint32_t)(x - g->boff[k]);
}

/* -------------------------------------------------------------------- forest */

typedef struct { int64_t id; uint32_t row; } id_row;

static int cmp_id_row(const void* a, const void* b) {
    const id_row* x = (const id_row*)a; const id_row* y = (const id_row*)b;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

static int64_t find_row(const id_row* ix, uint32_t n, int64_t id) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (ix[mid].id < id) lo = mid + 1; else hi = mid;
    }
    return (lo < n && ix[lo].id == id) ? (int64_t)ix[lo].row : -1;
}

void orc_forest_free(orc_forest* f) {
    if (!f) return;
    free(f->cell_id); free(f->parent); free(f->species); free(f->birth); free(f->leaf_lo);
    free(f->leaf_hi); free(f->sample_begin); free(f->sample_leaves);
    free(f);
}

/* SamplesForest node table (src/forest.hpp:29-67) in DFS pre-order: roots and
 * children by ascending cell id; leaves are the childless nodes. */
int orc_forest_flatten(const psx_forest_desc* d, orc_forest** out) {
    uint32_t n = d->n_nodes;
    if (n == 0) return fail("the forest has no nodes");
    id_row* ix = (id_row*)malloc(sizeof(id_row) * n);
    for (uint32_t i = 0; i < n; ++i) { ix[i].id = d->cell_id[i]; ix[i].row = i; }
    qsort(ix, n, sizeof(id_row), cmp_id_row);
    for (uint32_t i = 1; i < n; ++i)
        if (ix[i].id == ix[i - 1].id) {
            const long long dup = (long long)ix[i].id;
            free(ix);
            return fail("duplicated cell id %lld", dup);
        }
    uint32_t* prow = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* nchild = (uint32_t*)calloc(n + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        if (d->parent[i] < 0) { prow[i] = PSX_NO_NODE; continue; }
        int64_t r = find_row(ix, n, d->parent[i]);
        if (r < 0) { free(ix); free(prow); free(nchild); return fail("unknown ancestor %lld", (long long)d->parent[i]); }
        prow[i] = (uint32_t)r;
        nchild[r + 1]++;
    }
    for (uint32_t i = 0; i < n; ++i) nchild[i + 1] += nchild[i];
    uint32_t* child = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t* fill = (uint32_t*)calloc(n, sizeof(uint32_t));
    /* visiting rows in cell-id order leaves every child list sorted by id */
    for (uint32_t k = 0; k < n; ++k) {
        uint32_t i = ix[k].row;
        if (prow[i] != PSX_NO_NODE) child[nchild[prow[i]] + fill[prow[i]]++] = i;
    }
    free(fill);
    orc_forest* f = (orc_forest*)calloc(1, sizeof(orc_forest));
    f->n_nodes = n;
    f->cell_id = (int64_t*)malloc(sizeof(int64_t) * n);
    f->parent = (uint32_t*)malloc(sizeof(uint32_t) * n);
    f->species = (uint32_t*)malloc(sizeof(uint32_t) * n);
    f->birth = (double*)malloc(sizeof(double) * n);
    f->leaf_lo = (uint32_t*)malloc(sizeof(uint32_t) * n);
    f->leaf_hi = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* dfs_of_row = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* row_of_dfs = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* stack = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t top = 0, next = 0;
    for (int64_t k = (int64_t)n - 1; k >= 0; --k) {
        uint32_t i = ix[k].row;
        if (prow[i] == PSX_NO_NODE) stack[top++] = i;
    }
    while (top) {
        uint32_t i = stack[--top];
        dfs_of_row[i] = next;
        row_of_dfs[next++] = i;
        for (uint32_t c = nchild[i + 1]; c > nchild[i]; --c) stack[top++] = child[c - 1];
    }
    free(stack); free(ix);
    if (next != n) {
        free(prow); free(nchild); free(child); free(dfs_of_row); free(row_of_dfs);
        orc_forest_free(f);
        return fail("the ancestor relation has a cycle");
    }
    /* subtree sizes by reverse pre-order, then leaf prefix counts */
    uint32_t* size = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t v = 0; v < n; ++v) size[v] = 1;
    for (int64_t v = (int64_t)n - 1; v >= 0; --v) {
        uint32_t i = row_of_dfs[v];
        if (prow[i] != PSX_NO_NODE) size[dfs_of_row[prow[i]]] += size[v];
    }
    uint32_t* leaves_before = (uint32_t*)malloc(sizeof(uint32_t) * (n + 1));
    leaves_before[0] = 0;
    for (uint32_t v = 0; v < n; ++v) {
        uint32_t i = row_of_dfs[v];
        leaves_before[v + 1] = leaves_before[v] + (nchild[i + 1] == nchild[i] ? 1u : 0u);
    }
    f->n_leaves = leaves_before[n];
    f->n_samples = d->n_samples;
    f->sample_begin = (uint32_t*)calloc(d->n_samples + 1, sizeof(uint32_t));
    f->sample_leaves = (uint32_t*)malloc(sizeof(uint32_t) * (f->n_leaves ? f->n_leaves : 1));
    int rc = 0;
    for (uint32_t v = 0; v < n && !rc; ++v) {
        uint32_t i = row_of_dfs[v];
        f->cell_id[v] = d->cell_id[i];
        f->parent[v] = prow[i] == PSX_NO_NODE ? PSX_NO_NODE : dfs_of_row[prow[i]];
        f->species[v] = d->species[i];
        f->birth[v] = d->birth_time[i];
        f->leaf_lo[v] = leaves_before[v];
        f->leaf_hi[v] = leaves_before[v + size[v]];
        int leaf = nchild[i + 1] == nchild[i];
        int32_t s = d->sample[i];
        if (s >= 0 && (!leaf || (uint32_t)s >= d->n_samples))
            rc = fail("cell %lld: invalid sample %d", (long long)d->cell_id[i], s);
        else if (s >= 0)
            f->sample_begin[s + 1]++;
    }
    for (uint32_t s = 0; s < d->n_samples; ++s) f->sample_begin[s + 1] += f->sample_begin[s];
    uint32_t* fillc = (uint32_t*)calloc(d->n_samples + 1, sizeof(uint32_t));
    for (uint32_t v = 0; v < n && !rc; ++v) {
        int32_t s = d->sample[row_of_dfs[v]];
        if (s >= 0) f->sample_leaves[f->sample_begin[s] + fillc[s]++] = leaves_before[v];
    }
    free(fillc); free(size); free(leaves_before); free(prow); free(nchild); free(child);
    free(dfs_of_row); free(row_of_dfs);
    if (rc) { orc_forest_free(f); return rc; }
    *out = f;
    return 0;
}

/* --------------------------------------------------------------- SID order */

static uint64_t sid_key(const psx_mutation* r) {
    uint64_t payload = r->kind == PSX_SNV ? r->alt : (r->kind == PSX_INS ? r->seq : 0u);
    return ((uint64_t)r->chr << 58) | ((uint64_t)r->pos << 26) | ((uint64_t)r->kind << 24) |
           ((uint64_t)(r->len & 15u) << 20) | (payload & 0xFFFFFu);
}

typedef struct { uint64_t key; uint64_t id; } key_id;

static int cmp_key_id(const void* a, const void* b) {
    const key_id* x = (const key_id*)a; const key_id* y = (const key_id*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

int orc_sort_table(psx_mutation* m, uint64_t n) {
    key_id* k = (key_id*)malloc(sizeof(key_id) * (n ? n : 1));
    for (uint64_t i = 0; i < n; ++i) { k[i].key = sid_key(&m[i]); k[i].id = i; }
    qsort(k, n, sizeof(key_id), cmp_key_id);
    psx_mutation* tmp = (psx_mutation*)malloc(sizeof(psx_mutation) * (n ? n : 1));
    for (uint64_t i = 0; i < n; ++i) {
        memcpy(&tmp[i], &m[k[i].id], sizeof(psx_mutation));
        tmp[i].sid_head = (i == 0 || k[i].key != k[i - 1].key) ? 1 : 0;
    }
    memcpy(m, tmp, sizeof(psx_mutation) * n);
    free(tmp); free(k);
    return 0;
}

/* ------------------------------------------------------- SBS context index */

/* The context index of the reference engine (build_contex_index,
 * src/mutation_engine.cpp:213-261, a RACES ContextIndex built with
 * context_sampling and the driver loci left out), restated for this model:
 * a site is a position q with 1 <= q < len - 1; its context is the
 * pyrimidine-centred trinucleotide, read on the strand where b[q] is C or T,
 *     ctx = (mid == T) << 4 | five << 2 | three        (32 contexts),
 * so SBS class cls (sub = cls / 16, five, three) samples context
 * (sub >= 3) << 4 | (cls & 15). The index keeps, per context, every
 * sampling-th site in genome order from the first (ranks 0, s, 2s, ...) as a
 * 32-bit global position; the driver loci are skipped when a site is drawn
 * (same law as leaving them out). RACES's own sampling rule is not in the
 * reference: this one is a restatement, not a replay. */

typedef struct sbs_index {
    uint64_t count[32];   /* sites per context */
    uint64_t kept[32];    /* ceil(count / sampling) */
    uint64_t off[33];     /* list ranges per context */
    uint32_t* list;       /* global positions, ascending per context */
} sbs_index;

static uint32_t sbs_context(uint32_t b5, uint32_t m, uint32_t b3) {
    if (m == 1u || m == 3u) return ((m == 3u ? 1u : 0u) << 4) | (b5 << 2) | b3;
    return ((m == 0u ? 1u : 0u) << 4) | ((3u - b3) << 2) | (3u - b5);
}

static int sbs_index_build(const genome_view* g, uint32_t sampling, sbs_index* ix) {
    memset(ix, 0, sizeof *ix);
    if (sampling == 0) sampling = 1;
    if (g->boff[g->n_chr] > 0xFFFFFFFFull)
        return fail("the reference is too long for the context index (at most 2^32 - 1 bases)");
    for (uint32_t c = 0; c < g->n_chr; ++c)
        for (uint32_t q = 1; q + 1 < g->len[c]; ++q)
            ix->count[sbs_context(base_at(g, c, q - 1), base_at(g, c, q), base_at(g, c, q + 1))]++;
    uint64_t total = 0;
    for (uint32_t k = 0; k < 32; ++k) {
        ix->kept[k] = (ix->count[k] + sampling - 1) / sampling;
        ix->off[k] = total;
        total += ix->kept[k];
    }
    ix->off[32] = total;
    ix->list = (uint32_t*)malloc(sizeof(uint32_t) * (total ? total : 1));
    uint64_t rank[32] = {0};
    for (uint32_t c = 0; c < g->n_chr; ++c)
        for (uint32_t q = 1; q + 1 < g->len[c]; ++q) {
            const uint32_t k = sbs_context(base_at(g, c, q - 1), base_at(g, c, q), base_at(g, c, q + 1));
            if (rank[k] % sampling == 0) ix->list[ix->off[k] + rank[k] / sampling] = (uint32_t)(g->boff[c] + q);
            rank[k]++;
        }
    return 0;
}

int orc_sbs_index(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, uint32_t sampling,
                  uint64_t* counts, uint64_t* kept, uint32_t** list) {
    if (n_chr == 0 || n_chr > PSX_MAX_CHROMOSOMES) return fail("invalid number of chromosomes");
    genome_view g;
    gv_init(&g, n_chr, len, packed);
    sbs_index ix;
    int rc = sbs_index_build(&g, sampling, &ix);
    if (rc) { free(ix.list); return rc; }
    memcpy(counts, ix.count, sizeof ix.count);
    memcpy(kept, ix.kept, sizeof ix.kept);
    *list = ix.list;
    return 0;
}

/* ------------------------------------------------------ indel (ID83) classes */

/* COSMIC ID83 classes over the reference (the repeated-sequence index of the
 * reference engine, src/mutation_engine.cpp:289-359, restated for this
 * model). For a position q of a chromosome b[0..len) (the anchor base of the
 * indel is b[q - 1]) and a unit length u in 1..5, U = b[q..q+u) and q is a
 * run start when b[q-u..q) != U (q >= u, q + u <= len); rep(q, u) counts the
 * copies of U from q on. Sites of a class (a payload selects the inserted
 * unit of the insertions next to 0 copies):
 *   deletion, u = 1, base class B, n copies (1..6+): run start, B(b[q]) = B,
 *       rep = n;
 *   deletion, u >= 2, repeats r (0..5+): run start, rep = r + 1, and for
 *       r = 0 no microhomology;
 *   deletion with microhomology h, u >= 2: run start, rep = 1, and the first
 *       h bases of U equal the h bases after it (h < u);
 *   insertion, u = 1, base class B, n copies (1..5+): run start, B(b[q]) = B,
 *       rep = n; the inserted base is b[q];
 *   insertion, u >= 2, n copies (1..5+): run start, rep = n; inserted U;
 *   insertion next to 0 copies: the inserted unit s (u = 1: one of the two
 *       bases of class B, payload bit; u >= 2: any u-mer, payload = its 2u
 *       bits) differs from b[q..q+u) and from b[q-u..q).
 * The deleted or inserted unit follows the anchor: record pos = q - 1. B is
 * the pyrimidine of the base's pair (C for C/G, T for T/A). */

enum { ID_DEL1 = 0, ID_INS1 = 12, ID_DELR = 24, ID_INSR = 48, ID_MH = 72 };
#define ID_DENSE_TRIES 1024u  /* a class whose rejection needs more tries on average is listed */

static const uint32_t ID_MH_BASE[6] = {0, 0, 72, 73, 75, 78}; /* first MH class of u = 2..5 */
static const uint32_t ID_MH_MAX[6] = {0, 0, 1, 2, 3, 5};

typedef struct id_class {
    uint32_t del, u, b, n, h, zero; /* zero: insertion next to 0 copies */
} id_class;

static id_class id_decode(uint32_t c) {
    id_class k;
    memset(&k, 0, sizeof k);
    if (c < ID_INS1) { k.del = 1; k.u = 1; k.b = c / 6; k.n = c % 6 + 1; }
    else if (c < ID_DELR) { k.u = 1; k.b = (c - ID_INS1) / 6; k.n = (c - ID_INS1) % 6; k.zero = k.n == 0; }
    else if (c < ID_INSR) { k.del = 1; k.u = 2 + (c - ID_DELR) / 6; k.n = (c - ID_DELR) % 6 + 1; }
    else if (c < ID_MH) { k.u = 2 + (c - ID_INSR) / 6; k.n = (c - ID_INSR) % 6; k.zero = k.n == 0; }
    else {
        k.del = 1;
        for (uint32_t u = 5; u >= 2; --u)
            if (c >= ID_MH_BASE[u]) { k.u = u; k.h = c - ID_MH_BASE[u] + 1; k.n = 1; break; }
    }
    return k;
}

/* payloads of a class: 2 (1 bp insertion next to 0 copies), 4^u (longer
 * insertions next to 0 copies), 1 otherwise */
static uint32_t id_payloads(uint32_t c) {
    id_class k = id_decode(c);
    return k.zero ? (k.u == 1 ? 2u : 1u << (2 * k.u)) : 1u;
}

static uint32_t pyr_class(uint32_t base) { return (base == 1u || base == 2u) ? 0u : 1u; } /* C/G -> 0, T/A -> 1 */

static int units_equal(const genome_view* g, uint32_t c, uint32_t a, uint32_t b, uint32_t u) {
    for (uint32_t i = 0; i < u; ++i)
        if (base_at(g, c, a + i) != base_at(g, c, b + i)) return 0;
    return 1;
}

static uint32_t id_rep(const genome_view* g, uint32_t c, uint32_t q, uint32_t u, uint32_t cap) {
    uint32_t n = 1;
    while (n < cap && (uint64_t)q + (uint64_t)(n + 1) * u <= g->len[c] && units_equal(g, c, q, q + n * u, u)) ++n;
    return n;
}

static uint32_t id_mh(const genome_view* g, uint32_t c, uint32_t q, uint32_t u) {
    uint32_t h = 0;
    while (h + 1 < u && (uint64_t)q + u + h < g->len[c] && base_at(g, c, q + h) == base_at(g, c, q + u + h)) ++h;
    return h;
}

static uint32_t unit_at(const genome_view* g, uint32_t c, uint32_t q, uint32_t u) {
    uint32_t s = 0;
    for (uint32_t i = 0; i < u; ++i) s |= base_at(g, c, q + i) << (2 * i);
    return s;
}

/* the inserted unit of a class-zero payload */
static uint32_t id_zero_unit(uint32_t u, uint32_t b, uint32_t payload) {
    if (u == 1) return b == 0 ? (payload & 1u ? 2u : 1u) : (payload & 1u ? 0u : 3u);
    return payload;
}

int orc_id_is_site(const genome_view* g, uint32_t c, uint32_t q, uint32_t payload, uint32_t cls) {
    id_class k = id_decode(cls);
    const uint32_t u = k.u;
    if (q < u || (uint64_t)q + u > g->len[c]) return 0;
    if (k.zero) {
        if (payload >= id_payloads(cls)) return 0;
        uint32_t s = id_zero_unit(u, k.b, payload);
        return s != unit_at(g, c, q, u) && s != unit_at(g, c, q - u, u);
    }
    if (units_equal(g, c, q - u, q, u)) return 0; /* not a run start */
    if (u == 1 && pyr_class(base_at(g, c, q)) != k.b) return 0;
    if (!k.del) return id_rep(g, c, q, u, 5) == k.n;
    uint32_t n = id_rep(g, c, q, u, 6);
    if (u == 1) return n == k.n;
    if (cls >= ID_MH) {
        uint32_t h = n == 1 ? id_mh(g, c, q, u) : 0;
        return h >= 1 && (h < ID_MH_MAX[u] ? h : ID_MH_MAX[u]) == k.h;
    }
    if (n != k.n) return 0;
    return n >= 2 || id_mh(g, c, q, u) == 0;
}

/* The classes with copies that position q is a site of (a deletion and an
 * insertion class per unit length at a run start): written to cls, their
 * number returned. */
static uint32_t id_sites_at(const genome_view* g, uint32_t c, uint32_t q, uint32_t cls[10]) {
    uint32_t n = 0;
    for (uint32_t u = 1; u <= 5; ++u) {
        if (q < u || (uint64_t)q + u > g->len[c]) continue;
        uint32_t cur = unit_at(g, c, q, u);
        if (cur == unit_at(g, c, q - u, u)) continue; /* not a run start */
        uint32_t n6 = id_rep(g, c, q, u, 6), n5 = n6 < 5 ? n6 : 5;
        if (u == 1) {
            uint32_t b = pyr_class(cur);
            cls[n++] = ID_DEL1 + b * 6 + n6 - 1;
            cls[n++] = ID_INS1 + b * 6 + n5;
        } else {
            if (n6 >= 2) cls[n++] = ID_DELR + (u - 2) * 6 + n6 - 1;
            else {
                uint32_t h = id_mh(g, c, q, u);
                cls[n++] = h ? ID_MH_BASE[u] + (h < ID_MH_MAX[u] ? h : ID_MH_MAX[u]) - 1 : ID_DELR + (u - 2) * 6;
            }
            cls[n++] = ID_INSR + (u - 2) * 6 + n5;
        }
    }
    return n;
}

/* Sites of the insertion classes next to 0 copies at q: payloads counted per
 * class (index: u = 1 of class C, u = 1 of class T, then u = 2..5). */
static void id_zero_counts(const genome_view* g, uint32_t c, uint32_t q, uint64_t z[6]) {
    for (uint32_t u = 1; u <= 5; ++u) {
        if (q < u || (uint64_t)q + u > g->len[c]) continue;
        uint32_t cur = unit_at(g, c, q, u), prev = unit_at(g, c, q - u, u);
        if (u == 1) {
            for (uint32_t b = 0; b < 2; ++b)
                for (uint32_t bit = 0; bit < 2; ++bit) {
                    uint32_t s = id_zero_unit(1, b, bit);
                    if (s != cur && s != prev) z[b]++;
                }
        } else {
            z[u] += (1ull << (2 * u)) - (cur == prev ? 1u : 2u);
        }
    }
}

static const uint32_t ID_ZERO_CLASS[6] = {ID_INS1, ID_INS1 + 6, ID_INSR, ID_INSR + 6, ID_INSR + 12, ID_INSR + 18};

typedef struct id_index {
    uint64_t count[PSX_ID_CLASSES];
    uint8_t listed[PSX_ID_CLASSES];       /* sparse classes: sites listed */
    uint64_t off[PSX_ID_CLASSES + 1];     /* list ranges per class */
    uint64_t* list;                       /* global q << 10 | payload, ascending per class */
} id_index;

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static int id_index_build(const genome_view* g, id_index* ix) {
    memset(ix, 0, sizeof *ix);
    const uint64_t G = g->boff[g->n_chr];
    for (uint32_t c = 0; c < g->n_chr; ++c) {
        const int64_t len = g->len[c];
#pragma omp parallel
        {
            uint64_t cnt[PSX_ID_CLASSES] = {0}, z[6] = {0};
            uint32_t cls[10];
#pragma omp for schedule(static)
            for (int64_t q = 0; q < len; ++q) {
                uint32_t n = id_sites_at(g, c, (uint32_t)q, cls);
                for (uint32_t i = 0; i < n; ++i) cnt[cls[i]]++;
                id_zero_counts(g, c, (uint32_t)q, z);
            }
            for (uint32_t i = 0; i < 6; ++i) cnt[ID_ZERO_CLASS[i]] += z[i];
#pragma omp critical
            for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) ix->count[k] += cnt[k];
        }
    }
    uint64_t total = 0;
    for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) {
        ix->off[k] = total;
        unsigned __int128 need = (unsigned __int128)G * id_payloads(k);
        ix->listed[k] = ix->count[k] > 0 && (unsigned __int128)ix->count[k] * ID_DENSE_TRIES < need;
        if (ix->listed[k]) total += ix->count[k];
    }
    ix->off[PSX_ID_CLASSES] = total;
    ix->list = (uint64_t*)malloc(sizeof(uint64_t) * (total ? total : 1));
    uint64_t fill[PSX_ID_CLASSES];
    for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) fill[k] = ix->off[k];
    int zero_listed = 0;
    for (uint32_t i = 0; i < 6; ++i) zero_listed |= ix->listed[ID_ZERO_CLASS[i]];
    for (uint32_t c = 0; c < g->n_chr && total; ++c) {
        const int64_t len = g->len[c];
#pragma omp parallel for schedule(static)
        for (int64_t q = 0; q < len; ++q) {
            uint32_t cls[10];
            uint32_t n = id_sites_at(g, c, (uint32_t)q, cls);
            const uint64_t gq = g->boff[c] + (uint64_t)q;
            for (uint32_t i = 0; i < n; ++i) {
                if (!ix->listed[cls[i]]) continue;
                uint64_t slot;
#pragma omp atomic capture
                slot = fill[cls[i]]++;
                ix->list[slot] = gq << 10;
            }
            if (!zero_listed) continue;
            for (uint32_t i = 0; i < 6; ++i) {
                const uint32_t k = ID_ZERO_CLASS[i];
                if (!ix->listed[k]) continue;
                const uint32_t np = id_payloads(k);
                for (uint32_t pl = 0; pl < np; ++pl) {
                    if (!orc_id_is_site(g, c, (uint32_t)q, pl, k)) continue;
                    uint64_t slot;
#pragma omp atomic capture
                    slot = fill[k]++;
                    ix->list[slot] = (gq << 10) | pl;
                }
            }
        }
    }
    for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) {
        if (fill[k] != ix->off[k + 1])
            return fail("indel index: class %u lists %llu sites, counted %llu", k,
                        (unsigned long long)(fill[k] - ix->off[k]), (unsigned long long)ix->count[k]);
        qsort(ix->list + ix->off[k], ix->off[k + 1] - ix->off[k], sizeof(uint64_t), cmp_u64);
    }
    return 0;
}

int orc_id_index(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, uint64_t* counts, uint8_t* listed,
                 uint64_t* off, uint64_t** list) {
    if (n_chr == 0 || n_chr > PSX_MAX_CHROMOSOMES) return fail("invalid number of chromosomes");
    genome_view g;
    gv_init(&g, n_chr, len, packed);
    id_index ix;
    int rc = id_index_build(&g, &ix);
    if (rc) { free(ix.list); return rc; }
    memcpy(counts, ix.count, sizeof ix.count);
    memcpy(listed, ix.listed, sizeof ix.listed);
    memcpy(off, ix.off, sizeof ix.off);
    *list = ix.list;
    return 0;
}

int orc_id_site(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, uint32_t chr, uint32_t q,
                uint32_t payload, uint32_t cls) {
    genome_view g;
    gv_init(&g, n_chr, len, packed);
    if (chr >= n_chr || cls >= PSX_ID_CLASSES) return 0;
    return orc_id_is_site(&g, chr, q, payload, cls);
}

/* ----------------------------------------------------------------- placement */

typedef struct { uint32_t k0, n; uint64_t* T; } table;

enum { SLOT_DRIVER, SLOT_GERM, SLOT_PRE_SNV, SLOT_PRE_INDEL, SLOT_SNV, SLOT_INDEL };

typedef struct slot {
    uint8_t type;
    uint32_t node;   /* DFS index or PSX_NO_NODE */
    uint32_t index;  /* driver / germline index, or rank within the node */
    uint32_t attempt;
} slot;

typedef struct place_ctx {
    genome_view g;
    const orc_forest* f;
    const psx_engine_params* p;
    uint64_t seed, G;
    table* expo;     /* per epoch, over signatures */
    table* cls;      /* per signature, over the 96 classes whose context has an indexed site */
    sbs_index sbs;   /* SBS context index at the engine's sampling */
    uint64_t* drv_pos;  /* driver loci (global positions), ascending, unique */
    uint64_t n_drv_pos;
    table len_tab;   /* indel lengths, value = 1 + index */
    int use_id;      /* indels from ID signatures */
    id_index idx;    /* ID classes of the reference */
    table* id_expo;  /* per epoch, over ID signatures */
    table* id_cls;   /* per ID signature, over the classes with a site */
} place_ctx;

static const uint32_t SUB_TARGET[6] = {0, 2, 3, 0, 1, 2}; /* C>A C>G C>T T>A T>C T>G */

static uint32_t epoch_of(const psx_engine_params* p, double t) {
    uint32_t e = 0;
    for (uint32_t k = 0; k < p->n_epochs; ++k)
        if (p->epoch_start[k] <= t) e = k;
    return e;
}

/* SBS passenger: signature from the timed exposure at the cell's birth
 * (MutationEngine$add_exposure, src/mutation_engine.cpp:734-739), class from
 * the signature, site uniform over the context index's sites of the class's
 * context other than the driver loci (src/mutation_engine.cpp:244-261). */
static int driver_locus(const place_ctx* x, uint64_t gp) {
    uint64_t lo = 0, hi = x->n_drv_pos;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (x->drv_pos[mid] < gp) lo = mid + 1; else hi = mid;
    }
    return lo < x->n_drv_pos && x->drv_pos[lo] == gp;
}

static int draw_snv(const place_ctx* x, uint32_t v, uint32_t k, uint32_t a, int fixed_sig,
                    psx_mutation* r) {
    uint32_t w[4];
    philox4(x->seed, k, v, a, (uint32_t)P_SNV << 24, w);
    uint32_t sig;
    if (fixed_sig >= 0) sig = (uint32_t)fixed_sig;
    else {
        const table* e = &x->expo[epoch_of(x->p, x->f->birth[v])];
        sig = pick(e->T, e->n, w[0]);
    }
    const table* ct = &x->cls[sig];
    uint32_t cls = pick(ct->T, ct->n, w[1]);
    uint32_t allele = mulhi32(w[2], x->p->ploidy);
    uint32_t sub = cls / 16, tgt = SUB_TARGET[sub];
    uint32_t ctx = ((sub >= 3 ? 1u : 0u) << 4) | (cls & 15u);
    const uint64_t n = x->sbs.kept[ctx];
    for (uint32_t t = 0; t < LIST_TRIES && n; ++t) {
        uint32_t ws[4];
        philox4(x->seed, k, v, a, ((uint32_t)P_SITE << 24) | t, ws);
        const uint32_t gp = x->sbs.list[x->sbs.off[ctx] + mulhi64(ws[0], ws[1], n)];
        if (driver_locus(x, gp)) continue;
        uint32_t c, pos;
        locate(&x->g, gp, &c, &pos);
        uint32_t m = base_at(&x->g, c, pos);
        r->chr = c; r->pos = pos; r->kind = PSX_SNV; r->ref = (uint8_t)m;
        r->alt = (uint8_t)((m == 1u || m == 3u) ? tgt : 3u - tgt);
        r->len = 0; r->seq = 0; r->allele = (uint8_t)allele; r->cause = (uint16_t)sig;
        return 0;
    }
    return fail("no reference site matches a drawn SBS or ID class");
}

/* Indel: deletion or insertion (top bit), geometric length, uniform position
 * with room for the deleted bases, inserted bases from a second draw. */
static int draw_indel(const place_ctx* x, uint32_t v, uint32_t k, uint32_t a, psx_mutation* r) {
    for (uint32_t t = 0; t < INDEL_TRIES; ++t) {
        uint32_t w[4];
        philox4(x->seed, k, v, a, ((uint32_t)P_INDEL << 24) | t, w);
        uint32_t kind = (w[0] >> 31) ? PSX_INS : PSX_DEL;
        uint32_t allele = mulhi32(w[0] << 1, x->p->ploidy);
        uint32_t l = 1 + pick(x->len_tab.T, x->len_tab.n, w[1]);
        uint32_t c, pos;
        locate(&x->g, mulhi64(w[2], w[3], x->G), &c, &pos);
        if ((uint64_t)pos + l >= x->g.len[c]) continue;
        uint32_t seq = 0;
        if (kind == PSX_DEL) {
            for (uint32_t i = 0; i < l; ++i) seq |= base_at(&x->g, c, pos + 1 + i) << (2 * i);
        } else {
            uint32_t s[4];
            philox4(x->seed, k, v, a, ((uint32_t)P_INDEL << 24) | 0x400000u | t, s);
            seq = s[0] & ((1u << (2 * l)) - 1u);
        }
        r->chr = c; r->pos = pos; r->kind = (uint8_t)kind; r->ref = (uint8_t)base_at(&x->g, c, pos);
        r->alt = 0; r->len = (uint8_t)l; r->seq = seq; r->allele = (uint8_t)allele;
        r->cause = PSX_NO_CAUSE;
        return 0;
    }
    return fail("no room for an indel");
}

/* Indel from ID signatures: signature from the ID exposure at the cell's
 * birth, ID class from the signature (over the classes with a site in the
 * reference), allele uniform, site uniform among the class's sites: by
 * rejection over (position, payload) for classes dense enough, by a uniform
 * index into the class's sorted site list otherwise. */
static int draw_indel_id(const place_ctx* x, uint32_t v, uint32_t k, uint32_t a, int fixed_sig, psx_mutation* r) {
    uint32_t w[4];
    philox4(x->seed, k, v, a, (uint32_t)P_IDCLS << 24, w);
    uint32_t sig;
    if (fixed_sig >= 0) sig = (uint32_t)fixed_sig;
    else {
        const table* e = &x->id_expo[epoch_of(x->p, x->f->birth[v])];
        sig = pick(e->T, e->n, w[0]);
    }
    const table* ct = &x->id_cls[sig];
    const uint32_t cls = pick(ct->T, ct->n, w[1]);
    const uint32_t allele = mulhi32(w[2], x->p->ploidy);
    uint64_t gq = 0;
    uint32_t payload = 0, c = 0, q = 0;
    if (x->idx.listed[cls]) {
        uint32_t ws[4];
        philox4(x->seed, k, v, a, (uint32_t)P_IDSITE << 24, ws);
        const uint64_t e = x->idx.list[x->idx.off[cls] + mulhi64(ws[0], ws[1], x->idx.count[cls])];
        gq = e >> 10;
        payload = (uint32_t)(e & 1023u);
        locate(&x->g, gq, &c, &q);
    } else {
        const uint32_t np = id_payloads(cls);
        uint32_t t;
        for (t = 0; t < SITE_TRIES; ++t) {
            uint32_t ws[4];
            philox4(x->seed, k, v, a, ((uint32_t)P_IDSITE << 24) | t, ws);
            gq = mulhi64(ws[0], ws[1], x->G);
            payload = mulhi32(ws[2], np);
            locate(&x->g, gq, &c, &q);
            if (orc_id_is_site(&x->g, c, q, payload, cls)) break;
        }
        if (t == SITE_TRIES) return fail("no reference site matches ID class %u", cls);
    }
    const id_class kk = id_decode(cls);
    r->chr = c; r->pos = q - 1; r->kind = kk.del ? PSX_DEL : PSX_INS;
    r->ref = (uint8_t)base_at(&x->g, c, q - 1); r->alt = 0; r->len = (uint8_t)kk.u;
    r->seq = kk.zero ? id_zero_unit(kk.u, kk.b, payload) : unit_at(&x->g, c, q, kk.u);
    r->allele = (uint8_t)allele; r->cause = (uint16_t)(x->p->n_sbs + sig);
    return 0;
}

static int draw_slot(const place_ctx* x, const slot* s, psx_mutation* r) {
    const orc_forest* f = x->f;
    const psx_engine_params* p = x->p;
    memset(r, 0, sizeof *r);
    int rc = 0;
    switch (s->type) {
    case SLOT_DRIVER: {
        const psx_driver* d = &p->drivers[s->index];
        uint32_t allele;
        if (d->allele >= 0) allele = (uint32_t)d->allele;
        else {
            uint32_t w[4];
            philox4(x->seed, s->index, s->node, 0, (uint32_t)P_DRIVER << 24, w);
            allele = mulhi32(w[0], p->ploidy);
        }
        r->chr = d->chr; r->pos = d->pos; r->kind = PSX_SNV;
        r->ref = (uint8_t)base_at(&x->g, d->chr, d->pos); r->alt = d->alt;
        r->allele = (uint8_t)allele; r->cause = PSX_NO_CAUSE; r->nature = PSX_DRIVER;
        r->node = s->node; r->leaf_lo = f->leaf_lo[s->node]; r->leaf_hi = f->leaf_hi[s->node];
        return 0;
    }
    case SLOT_GERM: {
        const psx_germline* gm = &p->germline[s->index];
        r->chr = gm->chr; r->pos = gm->pos; r->kind = PSX_SNV;
        r->ref = (uint8_t)base_at(&x->g, gm->chr, gm->pos); r->alt = gm->alt;
        r->allele = gm->allele; r->cause = PSX_NO_CAUSE; r->nature = PSX_GERMINAL;
        r->node = PSX_NO_NODE; r->leaf_lo = 0; r->leaf_hi = f->n_leaves;
        return 0;
    }
    case SLOT_PRE_SNV:
        rc = draw_snv(x, PSX_NO_NODE, s->index, s->attempt, (int)p->preneoplastic_sbs, r);
        r->nature = PSX_PRENEOPLASTIC; r->node = PSX_NO_NODE; r->leaf_lo = 0; r->leaf_hi = f->n_leaves;
        return rc;
    case SLOT_PRE_INDEL:
        rc = x->use_id ? draw_indel_id(x, PSX_NO_NODE, s->index, s->attempt, (int)p->preneoplastic_id, r)
                       : draw_indel(x, PSX_NO_NODE, s->index, s->attempt, r);
        r->nature = PSX_PRENEOPLASTIC; r->node = PSX_NO_NODE; r->leaf_lo = 0; r->leaf_hi = f->n_leaves;
        return rc;
    case SLOT_SNV:
        rc = draw_snv(x, s->node, s->index, s->attempt, -1, r);
        break;
    default:
        rc = x->use_id ? draw_indel_id(x, s->node, s->index, s->attempt, -1, r)
                       : draw_indel(x, s->node, s->index, s->attempt, r);
        break;
    }
    r->nature = PSX_PASSENGER; r->node = s->node;
    r->leaf_lo = f->leaf_lo[s->node]; r->leaf_hi = f->leaf_hi[s->node];
    return rc;
}

typedef struct { uint32_t chr, start, end; uint64_t id; } footprint;

static int cmp_fp(const void* a, const void* b) {
    const footprint* x = (const footprint*)a; const footprint* y = (const footprint*)b;
    if (x->chr != y->chr) return x->chr < y->chr ? -1 : 1;
    if (x->start != y->start) return x->start < y->start ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

static int touches(const footprint* a, const footprint* b) {
    return a->chr == b->chr && (uint64_t)b->start <= (uint64_t)a->end + 1 &&
           (uint64_t)a->start <= (uint64_t)b->end + 1;
}

static void free_tables(place_ctx* x) {
    if (x->expo) for (uint32_t e = 0; e < x->p->n_epochs; ++e) free(x->expo[e].T);
    if (x->cls) for (uint32_t s = 0; s < x->p->n_sbs; ++s) free(x->cls[s].T);
    if (x->id_expo) for (uint32_t e = 0; e < x->p->n_epochs; ++e) free(x->id_expo[e].T);
    if (x->id_cls) for (uint32_t s = 0; s < x->p->n_id; ++s) free(x->id_cls[s].T);
    free(x->expo); free(x->cls); free(x->len_tab.T); free(x->id_expo); free(x->id_cls); free(x->idx.list);
    free(x->sbs.list); free(x->drv_pos);
}

/* ID signature tables: per epoch over the signatures, and per signature over
 * the 83 classes with every class that has no site in the reference given
 * probability 0 (a signature left without mass is an error). */
static int id_tables(place_ctx* x) {
    const psx_engine_params* p = x->p;
    int rc = id_index_build(&x->g, &x->idx);
    x->id_expo = (table*)calloc(p->n_epochs ? p->n_epochs : 1, sizeof(table));
    x->id_cls = (table*)calloc(p->n_id, sizeof(table));
    for (uint32_t e = 0; e < p->n_epochs && !rc; ++e) {
        x->id_expo[e].T = (uint64_t*)malloc(sizeof(uint64_t) * p->n_id);
        x->id_expo[e].n = p->n_id;
        rc = orc_cdf_table(p->id_exposure + (size_t)e * p->n_id, p->n_id, x->id_expo[e].T);
    }
    for (uint32_t s = 0; s < p->n_id && !rc; ++s) {
        double pmf[PSX_ID_CLASSES];
        double mass = 0.0;
        for (uint32_t k = 0; k < PSX_ID_CLASSES; ++k) {
            pmf[k] = x->idx.count[k] ? p->id[(size_t)s * PSX_ID_CLASSES + k] : 0.0;
            mass += pmf[k];
        }
        if (!(mass > 0.0)) return fail("ID signature %u has no class with a site in the reference", s);
        x->id_cls[s].T = (uint64_t*)malloc(sizeof(uint64_t) * PSX_ID_CLASSES);
        x->id_cls[s].n = PSX_ID_CLASSES;
        rc = orc_cdf_table(pmf, PSX_ID_CLASSES, x->id_cls[s].T);
    }
    return rc;
}

/* MutationEngine::place_mutations (src/mutation_engine.cpp:1085-1113): per
 * non-root edge Poisson(rate(species) * ploidy * G) SNVs and indels, the
 * preneoplastic SNVs/indels shared by every tumour cell, drivers on the edge
 * where their species first appears, germline in every cell. With the
 * infinite-sites model (src/mutation_engine.hpp:148-156) no two mutations
 * have touching footprints: the later-created one is redrawn. */
int orc_place_mutations(uint32_t n_chr, const uint32_t* len, const uint32_t* packed,
                        const orc_forest* f, const psx_engine_params* p, uint64_t seed,
                        psx_mutation** out, uint64_t* n_out) {
    if (n_chr == 0 || n_chr > PSX_MAX_CHROMOSOMES) return fail("invalid number of chromosomes");
    if (p->ploidy < 1 || p->ploidy > 255) return fail("invalid ploidy");
    if (p->indel_max_len < 1 || p->indel_max_len > 10) return fail("indel_max_len must be in [1, 10]");
    if (!(p->indel_len_p > 0.0 && p->indel_len_p <= 1.0)) return fail("invalid indel length parameter");
    for (uint32_t v = 0; v < f->n_nodes; ++v)
        if (f->species[v] >= p->n_species) return fail("node species %u has no rates", f->species[v]);
    place_ctx x;
    memset(&x, 0, sizeof x);
    gv_init(&x.g, n_chr, len, packed);
    x.f = f; x.p = p; x.seed = seed; x.G = x.g.boff[n_chr];
    int need_snv = p->n_preneoplastic_snv > 0;
    for (uint32_t s = 0; s < p->n_species; ++s) need_snv |= p->snv_rate[s] > 0.0;
    if (need_snv && (p->n_sbs == 0 || p->n_epochs == 0)) return fail("SNVs need a signature and an exposure");
    if (p->n_preneoplastic_snv > 0 && p->preneoplastic_sbs >= p->n_sbs) return fail("unknown preneoplastic signature");
    int rc = 0;
    x.expo = (table*)calloc(p->n_epochs ? p->n_epochs : 1, sizeof(table));
    x.cls = (table*)calloc(p->n_sbs ? p->n_sbs : 1, sizeof(table));
    for (uint32_t e = 0; e < p->n_epochs && !rc; ++e) {
        x.expo[e].T = (uint64_t*)malloc(sizeof(uint64_t) * p->n_sbs);
        x.expo[e].n = p->n_sbs;
        rc = orc_cdf_table(p->exposure + (size_t)e * p->n_sbs, p->n_sbs, x.expo[e].T);
    }
    /* SBS classes: the context index at the engine's sampling, every class whose context has no indexed site
     * given probability 0 (a signature left without mass is an error) */
    if (need_snv) rc = sbs_index_build(&x.g, p->context_sampling, &x.sbs);
    for (uint32_t s = 0; s < p->n_sbs && !rc && need_snv; ++s) {
        double pmf[96];
        double mass = 0.0;
        for (uint32_t k = 0; k < 96; ++k) {
            pmf[k] = x.sbs.kept[((k >= 48 ? 1u : 0u) << 4) | (k & 15u)] ? p->sbs[(size_t)s * 96 + k] : 0.0;
            mass += pmf[k];
        }
        if (!(mass > 0.0)) rc = fail("SBS signature %u has no class with a site in the context index", s);
        x.cls[s].T = (uint64_t*)malloc(sizeof(uint64_t) * 96);
        x.cls[s].n = 96;
        if (!rc) rc = orc_cdf_table(pmf, 96, x.cls[s].T);
    }
    x.drv_pos = (uint64_t*)malloc(sizeof(uint64_t) * (p->n_drivers ? p->n_drivers : 1));
    for (uint32_t d = 0; d < p->n_drivers && !rc; ++d)
        if (p->drivers[d].chr < n_chr) x.drv_pos[x.n_drv_pos++] = x.g.boff[p->drivers[d].chr] + p->drivers[d].pos;
    qsort(x.drv_pos, x.n_drv_pos, sizeof(uint64_t), cmp_u64);
    {
        uint64_t u = 0;
        for (uint64_t i = 0; i < x.n_drv_pos; ++i)
            if (u == 0 || x.drv_pos[u - 1] != x.drv_pos[i]) x.drv_pos[u++] = x.drv_pos[i];
        x.n_drv_pos = u;
    }
    if (p->n_id) {
        int need_id = p->n_preneoplastic_indel > 0;
        for (uint32_t s = 0; s < p->n_species; ++s) need_id |= p->indel_rate[s] > 0.0;
        if (!p->id || (need_id && (!p->id_exposure || p->n_epochs == 0)))
            rc = fail("indels need an ID signature and an exposure");
        else if (p->n_preneoplastic_indel > 0 && p->preneoplastic_id >= p->n_id)
            rc = fail("unknown preneoplastic ID signature");
        else if (need_id) {
            x.use_id = 1;
            rc = id_tables(&x);
        }
    }
    double lp[10];
    for (uint32_t l = 1; l <= p->indel_max_len; ++l)
        lp[l - 1] = pow(1.0 - p->indel_len_p, (double)(l - 1)) * p->indel_len_p;
    x.len_tab.T = (uint64_t*)malloc(sizeof(uint64_t) * 10);
    x.len_tab.n = p->indel_max_len;
    if (!rc) rc = orc_cdf_table(lp, p->indel_max_len, x.len_tab.T);
    table* ts = (table*)calloc(p->n_species, sizeof(table));
    table* ti = (table*)calloc(p->n_species, sizeof(table));
    double allelic = (double)((uint64_t)p->ploidy * x.G);
    for (uint32_t s = 0; s < p->n_species && !rc; ++s) {
        ts[s].T = (uint64_t*)malloc(sizeof(uint64_t) * TABLE_CAP);
        ti[s].T = (uint64_t*)malloc(sizeof(uint64_t) * TABLE_CAP);
        rc = orc_poisson_table(p->snv_rate[s] * allelic, &ts[s].k0, &ts[s].n, ts[s].T, TABLE_CAP);
        if (!rc) rc = orc_poisson_table(p->indel_rate[s] * allelic, &ti[s].k0, &ti[s].n, ti[s].T, TABLE_CAP);
    }
    for (uint32_t d = 0; d < p->n_drivers && !rc; ++d) {
        const psx_driver* dr = &p->drivers[d];
        if (dr->chr >= n_chr || dr->pos >= len[dr->chr] || dr->alt > 3 ||
            (dr->allele >= 0 && (uint32_t)dr->allele >= p->ploidy) ||
            dr->alt == base_at(&x.g, dr->chr, dr->pos))
            rc = fail("invalid driver %u", d);
    }
    for (uint32_t gi = 0; gi < p->n_germline && !rc; ++gi) {
        const psx_germline* gm = &p->germline[gi];
        if (gm->chr >= n_chr || gm->pos >= len[gm->chr] || gm->alt > 3 || gm->allele >= p->ploidy ||
            gm->alt == base_at(&x.g, gm->chr, gm->pos))
            rc = fail("invalid germline mutation %u", gi);
    }
    if (rc) {
        for (uint32_t s = 0; s < p->n_species; ++s) { free(ts[s].T); free(ti[s].T); }
        free(ts); free(ti); free_tables(&x);
        return rc;
    }
    /* creation order = mutation id */
    uint32_t* cs = (uint32_t*)calloc(f->n_nodes, sizeof(uint32_t));
    uint32_t* ci = (uint32_t*)calloc(f->n_nodes, sizeof(uint32_t));
    uint64_t total = (uint64_t)p->n_germline + p->n_preneoplastic_snv + p->n_preneoplastic_indel;
    uint64_t n_driver_slots = 0;
    for (uint32_t d = 0; d < p->n_drivers; ++d)
        for (uint32_t v = 0; v < f->n_nodes; ++v) {
            uint32_t par = f->parent[v];
            if (f->species[v] == p->drivers[d].species &&
                (par == PSX_NO_NODE || f->species[par] != p->drivers[d].species))
                ++n_driver_slots;
        }
    total += n_driver_slots;
    for (uint32_t v = 0; v < f->n_nodes; ++v) {
        if (f->parent[v] == PSX_NO_NODE) continue;
        uint32_t w[4];
        philox4(seed, v, 0, 0, (uint32_t)P_EDGE << 24, w);
        uint32_t s = f->species[v];
        cs[v] = ts[s].k0 + pick(ts[s].T, ts[s].n, w[0]);
        ci[v] = ti[s].k0 + pick(ti[s].T, ti[s].n, w[1]);
        total += (uint64_t)cs[v] + ci[v];
    }
    for (uint32_t s = 0; s < p->n_species; ++s) { free(ts[s].T); free(ti[s].T); }
    free(ts); free(ti);
    slot* sl = (slot*)malloc(sizeof(slot) * (total ? total : 1));
    uint64_t id = 0;
    for (uint32_t d = 0; d < p->n_drivers; ++d)
        for (uint32_t v = 0; v < f->n_nodes; ++v) {
            uint32_t par = f->parent[v];
            if (f->species[v] == p->drivers[d].species &&
                (par == PSX_NO_NODE || f->species[par] != p->drivers[d].species))
                sl[id++] = (slot){SLOT_DRIVER, v, d, 0};
        }
    for (uint32_t gi = 0; gi < p->n_germline; ++gi) sl[id++] = (slot){SLOT_GERM, PSX_NO_NODE, gi, 0};
    for (uint32_t k = 0; k < p->n_preneoplastic_snv; ++k) sl[id++] = (slot){SLOT_PRE_SNV, PSX_NO_NODE, k, 0};
    for (uint32_t k = 0; k < p->n_preneoplastic_indel; ++k) sl[id++] = (slot){SLOT_PRE_INDEL, PSX_NO_NODE, k, 0};
    for (uint32_t v = 0; v < f->n_nodes; ++v) {
        for (uint32_t k = 0; k < cs[v]; ++k) sl[id++] = (slot){SLOT_SNV, v, k, 0};
        for (uint32_t k = 0; k < ci[v]; ++k) sl[id++] = (slot){SLOT_INDEL, v, k, 0};
    }
    free(cs); free(ci);
    psx_mutation* m = (psx_mutation*)calloc(total ? total : 1, sizeof(psx_mutation));
    for (uint64_t i = 0; i < total && !rc; ++i) rc = draw_slot(&x, &sl[i], &m[i]);
    if (!rc && p->infinite_sites) {
        footprint* fp = (footprint*)malloc(sizeof(footprint) * (total ? total : 1));
        uint8_t* loser = (uint8_t*)malloc(total ? total : 1);
        int round;
        for (round = 0; round < IS_ROUNDS && !rc; ++round) {
            for (uint64_t i = 0; i < total; ++i) {
                fp[i].chr = m[i].chr; fp[i].start = m[i].pos;
                fp[i].end = m[i].pos + (m[i].kind == PSX_DEL ? m[i].len : 0u);
                fp[i].id = i;
            }
            qsort(fp, total, sizeof(footprint), cmp_fp);
            memset(loser, 0, total ? total : 1);
            uint64_t n_lose = 0;
            for (uint64_t j = 0; j < total; ++j) {
                uint64_t idj = fp[j].id;
                if (sl[idj].type == SLOT_DRIVER || sl[idj].type == SLOT_GERM) continue;
                int lose = 0;
                for (uint64_t i = j; i-- > 0 && !lose;) {
                    if (fp[i].chr != fp[j].chr || fp[i].start + 12u < fp[j].start) break;
                    if (fp[i].id < idj && touches(&fp[i], &fp[j])) lose = 1;
                }
                for (uint64_t i = j + 1; i < total && !lose; ++i) {
                    if (fp[i].chr != fp[j].chr || (uint64_t)fp[i].start > (uint64_t)fp[j].end + 1) break;
                    if (fp[i].id < idj && touches(&fp[i], &fp[j])) lose = 1;
                }
                if (lose) { loser[idj] = 1; ++n_lose; }
            }
            if (n_lose == 0) break;
            for (uint64_t i = 0; i < total && !rc; ++i)
                if (loser[i]) { sl[i].attempt++; rc = draw_slot(&x, &sl[i], &m[i]); }
        }
        if (!rc && round == IS_ROUNDS) rc = fail("infinite-sites placement did not converge");
        free(fp); free(loser);
    }
    free(sl);
    free_tables(&x);
    if (rc) { free(m); return rc; }
    orc_sort_table(m, total);
    *out = m;
    *n_out = total;
    return 0;
}

/* ------------------------------------------------------ copy-number drivers */

int orc_place_cna(uint32_t n_chr, const uint32_t* len, const orc_forest* f, const psx_engine_params* p,
                  uint64_t seed, psx_cna_event** out, uint32_t* n_out) {
    *out = NULL;
    *n_out = 0;
    if (p->n_cna && p->ploidy > PSX_MAX_CNA_PLOIDY) return fail("copy-number events need a small ploidy");
    uint64_t G = 0;
    for (uint32_t c = 0; c < n_chr; ++c) G += len[c];
    uint64_t cap = 0;
    for (uint32_t k = 0; k < p->n_cna; ++k) {
        const psx_cna* c = &p->cna[k];
        int wgd = c->kind == PSX_CNA_WGD;
        if (c->kind > PSX_CNA_WGD || c->species >= p->n_species ||
            (!wgd && (c->chr >= n_chr || c->len == 0 || c->start >= len[c->chr] || c->len > len[c->chr] - c->start ||
                      (c->allele >= p->ploidy && c->allele != PSX_CNA_RANDOM_ALLELE) ||
                      (c->kind == PSX_CNA_AMP && c->copies == 0))))
            return fail("invalid copy-number driver %u", k);
        for (uint32_t v = 0; v < f->n_nodes; ++v) {
            uint32_t par = f->parent[v];
            if (f->species[v] == c->species && (par == PSX_NO_NODE || f->species[par] != c->species)) ++cap;
        }
    }
    /* passenger CNAs (src/mutation_engine.cpp:441-487, 901-927): per non-root node Poisson(cna_rate(species) x
     * ploidy x G) events, the count from word z of the node's edge draw, each event a uniform region of the
     * catalogue on a uniform allele; an amplification adds one copy */
    int any = 0;
    for (uint32_t s = 0; s < p->n_species && p->cna_rate; ++s) {
        if (!(p->cna_rate[s] >= 0.0) || !isfinite(p->cna_rate[s])) return fail("CNA rates must be non-negative");
        any |= p->cna_rate[s] > 0.0;
    }
    uint32_t* npass = (uint32_t*)calloc(f->n_nodes ? f->n_nodes : 1, sizeof(uint32_t));
    if (any) {
        if (p->n_cna_catalogue == 0 || !p->cna_catalogue) { free(npass); return fail("passenger CNAs need a CNA catalogue"); }
        for (uint32_t k = 0; k < p->n_cna_catalogue; ++k) {
            const psx_cna_region* r = &p->cna_catalogue[k];
            if (r->chr >= n_chr || r->len == 0 || r->start >= len[r->chr] || r->len > len[r->chr] - r->start ||
                (r->kind != PSX_CNA_AMP && r->kind != PSX_CNA_DEL)) {
                free(npass);
                return fail("invalid passenger CNA catalogue region %u", k);
            }
        }
        double allelic = (double)((uint64_t)p->ploidy * G);
        uint64_t* T = (uint64_t*)malloc(sizeof(uint64_t) * TABLE_CAP);
        for (uint32_t v = 0; v < f->n_nodes; ++v) {
            if (f->parent[v] == PSX_NO_NODE) continue;
            double lam = p->cna_rate[f->species[v]] * allelic;
            if (lam == 0.0) continue;
            uint32_t k0, n;
            if (orc_poisson_table(lam, &k0, &n, T, TABLE_CAP)) { free(T); free(npass); return -1; }
            uint32_t w[4];
            philox4(seed, v, 0, 0, (uint32_t)P_EDGE << 24, w);
            npass[v] = k0 + pick(T, n, w[2]);
            cap += npass[v];
        }
        free(T);
    }
    psx_cna_event* ev = (psx_cna_event*)calloc(cap ? cap : 1, sizeof(psx_cna_event));
    uint32_t n = 0;
    /* nodes in DFS order; at one node, drivers in their given order, then the passengers */
    for (uint32_t v = 0; v < f->n_nodes; ++v) {
        uint32_t par = f->parent[v];
        for (uint32_t k = 0; k < p->n_cna; ++k) {
            const psx_cna* c = &p->cna[k];
            if (f->species[v] != c->species || (par != PSX_NO_NODE && f->species[par] == c->species)) continue;
            int wgd = c->kind == PSX_CNA_WGD;
            psx_cna_event* e = &ev[n++];
            e->node = v;
            e->leaf_lo = f->leaf_lo[v];
            e->leaf_hi = f->leaf_hi[v];
            e->chr = wgd ? PSX_ALL_CHROMOSOMES : c->chr;
            e->start = wgd ? 0 : c->start;
            e->end = wgd ? 0 : c->start + c->len;
            e->allele = wgd ? 0 : c->allele;
            if (!wgd && c->allele == PSX_CNA_RANDOM_ALLELE) {
                uint32_t w[4];
                philox4(seed, k, v, 1, (uint32_t)P_DRIVER << 24, w);
                e->allele = (uint8_t)mulhi32(w[0], p->ploidy);
            }
            e->kind = c->kind;
            e->copies = c->kind == PSX_CNA_AMP ? c->copies : 0;
        }
        for (uint32_t k = 0; k < npass[v]; ++k) {
            uint32_t w[4];
            philox4(seed, k, v, 0, (uint32_t)P_PCNA << 24, w);
            const psx_cna_region* r = &p->cna_catalogue[mulhi32(w[0], p->n_cna_catalogue)];
            psx_cna_event* e = &ev[n++];
            e->node = v;
            e->leaf_lo = f->leaf_lo[v];
            e->leaf_hi = f->leaf_hi[v];
            e->chr = r->chr;
            e->start = r->start;
            e->end = r->start + r->len;
            e->allele = (uint8_t)mulhi32(w[1], p->ploidy);
            e->kind = r->kind;
            e->copies = r->kind == PSX_CNA_AMP ? 1 : 0;
        }
    }
    free(npass);
    *out = ev;
    *n_out = n;
    return 0;
}

/* ---------------------------------------------------------------- sequencing */

typedef struct seq_ctx {
    const orc_forest* f;
    const psx_mutation* m;
    uint64_t n_mut;
    uint64_t chr_begin[PSX_MAX_CHROMOSOMES + 1];
    uint32_t ploidy, L, pe, preneo_in_normal;
    uint64_t seed;
    double e;
    uint32_t ins_k0, ins_n;
    uint64_t ins_T[4096];
    uint32_t err_N, err_k0, err_n;  /* bases per template, error-count table */
    uint64_t err_T[PSX_ERR_KCAP + 1];
    const psx_cna_event* ev;  /* copy-number events, DFS order */
    uint32_t n_ev;
    /* read materialisation (rows != NULL): the packed reference, the first
     * word of each chromosome, rw words per read, the first read of each
     * (output sample, chromosome) segment */
    const uint32_t* packed;
    uint64_t chr_woff[PSX_MAX_CHROMOSOMES];
    uint32_t* rows;
    uint32_t rw;
    /* a consumer of every read in place of the counting (orc_seq_walk) */
    orc_read_fn hook;
    void* hook_user;
} seq_ctx;

/* Copy numbers of the allele lineages of `leaf` at (c, x): one copy each,
 * then, in order, every event of the leaf's lineage covering x; returns the
 * total. */
static uint32_t copies(const seq_ctx* s, uint32_t c, uint32_t x, uint32_t leaf, uint32_t* cn) {
    for (uint32_t a = 0; a < s->ploidy; ++a) cn[a] = 1;
    for (uint32_t i = 0; i < s->n_ev; ++i) {
        const psx_cna_event* e = &s->ev[i];
        if (leaf < e->leaf_lo || leaf >= e->leaf_hi) continue;
        if (e->kind == PSX_CNA_WGD) {
            for (uint32_t a = 0; a < s->ploidy; ++a) cn[a] *= 2;
        } else if (e->chr == c && x >= e->start && x < e->end) {
            if (e->kind == PSX_CNA_AMP) cn[e->allele] += e->copies; else cn[e->allele] = 0;
        }
    }
    uint32_t t = 0;
    for (uint32_t a = 0; a < s->ploidy; ++a) t += cn[a];
    return t;
}

/* Events that act on chromosome c. */
static int chr_has_events(const seq_ctx* s, uint32_t c) {
    for (uint32_t i = 0; i < s->n_ev; ++i)
        if (s->ev[i].chr == c || s->ev[i].chr == PSX_ALL_CHROMOSOMES) return 1;
    return 0;
}

/* Largest copy number on chromosome c over every leaf and position (at
 * least the ploidy), by brute force over the leaves at event leaf-range
 * boundaries and the positions at event boundaries, between which the copy
 * numbers are constant. */
static uint32_t chr_kmax(const seq_ctx* s, uint32_t c, uint32_t len) {
    uint32_t best = s->ploidy, cn[PSX_MAX_CNA_PLOIDY];
    for (uint32_t i = 0; i <= s->n_ev; ++i)
        for (uint32_t side = 0; side < 2; ++side) {
            uint32_t leaf = i == s->n_ev ? 0 : (side ? s->ev[i].leaf_hi : s->ev[i].leaf_lo);
            if (leaf >= s->f->n_leaves) continue;
            for (uint32_t k = 0; k <= s->n_ev; ++k)
                for (uint32_t side2 = 0; side2 < 2; ++side2) {
                    uint32_t x = k == s->n_ev ? 0 : (s->ev[k].chr != c ? 0 : (side2 ? s->ev[k].end : s->ev[k].start));
                    if (x >= len) continue;
                    uint32_t t = copies(s, c, x, leaf, cn);
                    if (t > best) best = t;
                }
        }
    return best;
}

static int normal_carries(const seq_ctx* s, const psx_mutation* r) {
    return r->nature == PSX_GERMINAL || (r->nature == PSX_PRENEOPLASTIC && s->preneo_in_normal);
}

static int carried(const seq_ctx* s, const psx_mutation* r, uint32_t leaf, uint32_t allele) {
    if (r->allele != allele) return 0;
    if (leaf == s->f->n_leaves) return normal_carries(s, r);
    return r->leaf_lo <= leaf && leaf < r->leaf_hi;
}

/* Sequencing errors (independent per base: the base is replaced by one of
 * the three others, uniformly, with probability e), drawn per template r of
 * stratum j over its N = rpt L bases, base index b = mate L + offset. With
 * w_k = Philox(counter = (r, j, c2, ERROR << 24 | k)): the count K ~
 * Binomial(N, e) by inversion from word y of w_0; stream words are x, z, w of
 * w_0, then the four words of w_1, w_2, ...; each gives a base index
 * floor(w N / 2^32) and a substitution index floor((w & 0xFFFF) 3 / 2^16);
 * an index already taken is skipped; at most 4 PSX_ERR_KCAP stream words are
 * read. Returns the number of errors; b and sub receive them in draw order. */
static uint32_t template_errors(const seq_ctx* s, uint32_t r, uint32_t j, uint32_t c2, uint32_t* b, uint32_t* sub) {
    uint32_t w0[4], w[4];
    philox4(s->seed, r, j, c2, (uint32_t)P_ERROR << 24, w0);
    uint32_t K = s->err_k0 + pick(s->err_T, s->err_n, w0[1]);
    uint32_t cnt = 0;
    for (uint32_t t = 0; cnt < K && t < 4 * PSX_ERR_KCAP; ++t) {
        uint32_t word;
        if (t < 3) {
            word = w0[t == 0 ? 0 : t + 1];
        } else {
            if ((t - 3) % 4 == 0) philox4(s->seed, r, j, c2, ((uint32_t)P_ERROR << 24) | (1 + (t - 3) / 4), w);
            word = w[(t - 3) % 4];
        }
        uint32_t x = mulhi32(word, s->err_N);
        int dup = 0;
        for (uint32_t q = 0; q < cnt; ++q) dup |= b[q] == x;
        if (dup) continue;
        b[cnt] = x;
        sub[cnt] = ((word & 0xFFFFu) * 3u) >> 16;
        ++cnt;
    }
    return cnt;
}

/* One materialised read (the layout of psx_simulate_seq_reads): word 0 = start,
 * word 1 = leaf | allele << 26 | mate << 30 | (errors > 0) << 31, then base i
 * of the read in bits 2 (i % 16) of word 2 + i / 16: the reference base, the
 * first carried SNV at the position (table order), then the read's errors;
 * bits past the read end are 0. */
static void read_row(const seq_ctx* s, uint32_t c, uint32_t x, uint32_t leaf, uint32_t allele, uint32_t mate,
                     const uint32_t* off, const uint32_t* sub, uint32_t ne, uint32_t* row) {
    memset(row, 0, sizeof(uint32_t) * s->rw);
    row[0] = x;
    row[1] = leaf | (allele << 26) | (mate << 30) | ((ne > 0 ? 1u : 0u) << 31);
    uint32_t* bases = row + 2;
    for (uint32_t i = 0; i < s->L; ++i) {
        uint32_t p = x + i;
        uint32_t b = (s->packed[s->chr_woff[c] + p / 16] >> (2 * (p % 16))) & 3u;
        bases[i / 16] |= b << (2 * (i % 16));
    }
    uint64_t lo = s->chr_begin[c], hi = s->chr_begin[c + 1];
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (s->m[mid].pos < x) lo = mid + 1; else hi = mid;
    }
    uint32_t last = 0xFFFFFFFFu;
    for (uint64_t i = lo; i < s->chr_begin[c + 1] && s->m[i].pos < x + s->L; ++i) {
        const psx_mutation* r = &s->m[i];
        if (r->pos == last || r->kind != PSX_SNV || !carried(s, r, leaf, allele)) continue;
        uint32_t o = r->pos - x;
        bases[o / 16] = (bases[o / 16] & ~(3u << (2 * (o % 16)))) | ((uint32_t)r->alt << (2 * (o % 16)));
        last = r->pos;
    }
    for (uint32_t q = 0; q < ne; ++q) {
        uint32_t o = off[q], tb = (bases[o / 16] >> (2 * (o % 16))) & 3u;
        uint32_t nb = sub[q] < tb ? sub[q] : sub[q] + 1;
        bases[o / 16] = (bases[o / 16] & ~(3u << (2 * (o % 16)))) | (nb << (2 * (o % 16)));
    }
}

static void add32(int32_t* p, int32_t v) {
#ifdef _OPENMP
#pragma omp atomic
#endif
    *p += v;
}

/* One read [x, x + L) from source (leaf, allele): every mutation record at a
 * covered position gains depth; each SID head gains an alternative read when
 * the observed base equals its alternative (SNV; the true base is the first
 * carried SNV at the position, else the reference, then the read's error, if
 * any, substitutes it) or when the source carries the SID (indel). */
static void seq_read(const seq_ctx* s, uint32_t c, uint32_t x, uint32_t leaf, uint32_t allele,
                     const uint32_t* off, const uint32_t* sub, uint32_t ne, int32_t* alt, int32_t* depth) {
    uint64_t lo = s->chr_begin[c], hi = s->chr_begin[c + 1];
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (s->m[mid].pos < x) lo = mid + 1; else hi = mid;
    }
    uint64_t i = lo, end = s->chr_begin[c + 1];
    while (i < end && s->m[i].pos < x + s->L) {
        uint32_t pos = s->m[i].pos;
        uint64_t g0 = i;
        while (i < end && s->m[i].pos == pos) ++i;
        uint32_t tb = s->m[g0].ref;
        for (uint64_t q = g0; q < i; ++q)
            if (s->m[q].kind == PSX_SNV && carried(s, &s->m[q], leaf, allele)) { tb = s->m[q].alt; break; }
        uint32_t ob = tb;
        for (uint32_t q = 0; q < ne; ++q)
            if (off[q] == pos - x) ob = sub[q] < tb ? sub[q] : sub[q] + 1;
        for (uint64_t q = g0; q < i; ++q) {
            add32(&depth[q], 1);
            if (!s->m[q].sid_head) continue;
            if (s->m[q].kind == PSX_SNV) {
                if (ob == s->m[q].alt) add32(&alt[q], 1);
            } else {
                int car = 0;
                for (uint64_t q2 = q; q2 < i && (q2 == q || !s->m[q2].sid_head); ++q2)
                    car |= carried(s, &s->m[q2], leaf, allele);
                if (car) add32(&alt[q], 1);
            }
        }
    }
}

typedef struct { uint32_t so, c, j; } stratum_ref;

/* The ReadSimulator call of simulate_seq (src/seq_simulation.cpp:354-429)
 * restated as a count-only simulation. Per output sample and chromosome:
 * T templates (round(C * len / L), halved for paired ends) with starts
 * stratified over PSX_STRATUM-base strata of the W = len - L + 1 valid
 * starts (stratum j gets floor(T e_j / W) - floor(T b_j / W)); each template
 * draws its start, normal-vs-tumour source (probability 1 - purity), source
 * (cell, copy) uniformly, and for paired ends an insert size from get_bin_dist
 * (src/seq_simulation.cpp:431-451); the mate ends the fragment. */
int orc_simulate_seq(uint32_t n_chr, const uint32_t* len, const orc_forest* f,
                     const psx_mutation* m, uint64_t n_mut, uint32_t ploidy,
                     const psx_seq_params* p, uint64_t seed, int32_t* alt, int32_t* depth,
                     uint8_t* present, uint64_t* n_reads) {
    return orc_simulate_seq_cna(n_chr, len, f, m, n_mut, ploidy, NULL, 0, p, seed, alt, depth, present, n_reads);
}

/* simulate_seq; with rows != NULL also every read, materialised into rows
 * (capacity reads of psx_read_words words; no copy-number events). */
/* A template's source from its draw w (x its start): the normal genome with
 * probability 1 - purity (always for the normal sample), else a uniform
 * (cell, copy) of the sample. With copy-number events on the chromosome one
 * of K copy slots per cell is drawn and the template is kept only when the
 * slot is one of the source's copies at x (returns 0 otherwise). */
static int template_source(const seq_ctx* s, uint32_t c, int has_ev, uint32_t K, int is_normal, uint64_t thr_normal,
                           const uint32_t* sl, uint32_t ns, const uint32_t w[4], uint32_t x, uint32_t* leaf,
                           uint32_t* allele) {
    const uint32_t ploidy = s->ploidy;
    if (!has_ev) {
        if (is_normal || (uint64_t)w[1] < thr_normal) {
            *leaf = s->f->n_leaves;
            *allele = mulhi32(w[2], ploidy);
        } else {
            uint32_t X = mulhi32(w[2], ns * ploidy);
            *leaf = sl[X / ploidy];
            *allele = X % ploidy;
        }
        return 1;
    }
    if (is_normal || (uint64_t)w[1] < thr_normal) {
        uint32_t k = mulhi32(w[2], K);
        if (k >= ploidy) return 0;  /* not a copy of the normal genome */
        *leaf = s->f->n_leaves;
        *allele = k;
        return 1;
    }
    uint32_t X = mulhi32(w[2], ns * K);
    uint32_t k = X % K;
    *leaf = sl[X / K];
    uint32_t cn[PSX_MAX_CNA_PLOIDY];
    uint32_t tot = copies(s, c, x, *leaf, cn);
    if (k >= tot) return 0;  /* the cell has fewer copies at x */
    uint32_t a = 0, cum = cn[0];
    while (a + 1 < ploidy && k >= cum) cum += cn[++a];
    *allele = a;
    return 1;
}

static int seq_run(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, const orc_forest* f,
                   const psx_mutation* m, uint64_t n_mut, uint32_t ploidy, const psx_cna_event* ev, uint32_t n_ev,
                   const psx_seq_params* p, uint64_t seed, int32_t* alt, int32_t* depth, uint8_t* present,
                   uint32_t* rows, uint64_t capacity, orc_read_fn hook, void* hook_user, uint64_t* n_reads) {
    if (!(p->coverage >= 0.0) || !isfinite(p->coverage)) return fail("invalid coverage");
    if (p->read_size < 1) return fail("invalid read size");
    if (!(p->purity >= 0.0 && p->purity <= 1.0)) return fail("The purity must be in [0,1]");
    if (!(p->error_rate >= 0.0 && p->error_rate < 1.0))
        return fail("The parameter \"error_rate\" must be a positive real number.");
    if (ploidy < 1 || (uint64_t)f->n_leaves * ploidy >= ((uint64_t)1 << 32)) return fail("invalid ploidy");
    seq_ctx* s = (seq_ctx*)calloc(1, sizeof(seq_ctx));
    s->f = f; s->m = m; s->n_mut = n_mut; s->ploidy = ploidy; s->L = p->read_size;
    s->pe = p->insert_size_mean > 0; s->preneo_in_normal = p->preneoplastic_in_normal != 0;
    s->seed = seed; s->e = p->error_rate;
    s->ev = ev; s->n_ev = n_ev;
    s->packed = packed; s->rows = rows; s->rw = 2u + (s->L + 15u) / 16u;
    s->hook = hook; s->hook_user = hook_user;
    for (uint32_t c = 0, w = 0; c < n_chr && c < PSX_MAX_CHROMOSOMES; ++c) {
        s->chr_woff[c] = w;
        w += (len[c] + 15u) / 16u;
    }
    int rc = 0;
    if (n_ev && ploidy > PSX_MAX_CNA_PLOIDY) rc = fail("copy-number events need a small ploidy");
    if (s->pe) rc = insert_table(p->insert_size_mean, p->insert_size_stddev, &s->ins_k0, &s->ins_n, s->ins_T);
    if (!rc && s->e > 0.0) {
        s->err_N = (s->pe ? 2u : 1u) * s->L;
        if (!error_rate_ok(s->err_N, s->e)) rc = fail("error rate too high for %u-base reads", s->L);
        if (!rc) rc = error_table(s->err_N, s->e, &s->err_k0, &s->err_n, s->err_T);
    }
    /* chromosome set: sorted, unique (std::set semantics of get_relevant_chr_set) */
    uint8_t sel[PSX_MAX_CHROMOSOMES];
    memset(sel, 0, sizeof sel);
    if (p->n_chromosomes == 0) memset(sel, 1, n_chr);
    for (uint32_t k = 0; k < p->n_chromosomes && !rc; ++k) {
        if (p->chromosomes[k] >= n_chr) rc = fail("unknown chromosome %u", p->chromosomes[k]);
        else sel[p->chromosomes[k]] = 1;
    }
    uint32_t n_req = p->samples ? p->n_samples : f->n_samples;
    uint32_t* req = (uint32_t*)malloc(sizeof(uint32_t) * (n_req + 1));
    for (uint32_t k = 0; k < n_req && !rc; ++k) {
        req[k] = p->samples ? p->samples[k] : k;
        if (req[k] >= f->n_samples) rc = fail("unknown sample %u", req[k]);
        else if (f->sample_begin[req[k] + 1] == f->sample_begin[req[k]]) rc = fail("sample %u has no cells", req[k]);
    }
    for (uint64_t i = 1; i < n_mut && !rc; ++i)
        if (sid_key(&m[i]) < sid_key(&m[i - 1])) rc = fail("mutation table is not in SID order");
    if (rc) { free(req); free(s); return rc; }
    uint32_t n_out = n_req + (p->with_normal_sample ? 1u : 0u);
    {
        uint64_t i = 0;
        for (uint32_t c = 0; c <= n_chr; ++c) {
            while (i < n_mut && m[i].chr < c) ++i;
            s->chr_begin[c] = i;
        }
        for (uint32_t c = n_chr + 1; c <= PSX_MAX_CHROMOSOMES; ++c) s->chr_begin[c] = n_mut;
    }
    memset(alt, 0, sizeof(int32_t) * n_mut * n_out);
    memset(depth, 0, sizeof(int32_t) * n_mut * n_out);
    uint64_t thr_normal = (uint64_t)((1.0 - p->purity) * TWO32);
    /* present */
    for (uint32_t so = 0; so < n_out; ++so) {
        int is_normal = so == n_req;
        const uint32_t* sl = is_normal ? NULL : f->sample_leaves + f->sample_begin[req[so]];
        uint32_t ns = is_normal ? 0 : f->sample_begin[req[so] + 1] - f->sample_begin[req[so]];
        for (uint64_t i = 0; i < n_mut; ++i) {
            int pres = 0;
            if (is_normal || p->purity < 1.0) pres = normal_carries(s, &m[i]);
            if (!is_normal && !pres) {
                uint32_t lo = 0, hi = ns;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) / 2;
                    if (sl[mid] < m[i].leaf_lo) lo = mid + 1; else hi = mid;
                }
                pres = lo < ns && sl[lo] < m[i].leaf_hi;
            }
            present[(uint64_t)so * n_mut + i] = (uint8_t)pres;
        }
    }
    /* copy numbers: per chromosome, copy slots drawn per template */
    uint32_t kmax[PSX_MAX_CHROMOSOMES];
    uint8_t has_ev[PSX_MAX_CHROMOSOMES];
    for (uint32_t c = 0; c < n_chr; ++c) {
        has_ev[c] = (uint8_t)chr_has_events(s, c);
        kmax[c] = has_ev[c] ? chr_kmax(s, c, len[c]) : ploidy;
        if ((uint64_t)f->n_leaves * kmax[c] >= ((uint64_t)1 << 32)) {
            free(req); free(s);
            return fail("too many cell copies");
        }
    }
    /* strata in (output sample, chromosome, stratum) order */
    uint64_t n_strata = 0, reads = 0;
    for (uint32_t so = 0; so < n_out; ++so)
        for (uint32_t c = 0; c < n_chr; ++c)
            if (sel[c] && len[c] >= s->L) n_strata += (len[c] - s->L + 1 + PSX_STRATUM - 1) / PSX_STRATUM;
    stratum_ref* st = (stratum_ref*)malloc(sizeof(stratum_ref) * (n_strata ? n_strata : 1));
    uint64_t q = 0;
    for (uint32_t so = 0; so < n_out; ++so)
        for (uint32_t c = 0; c < n_chr; ++c) {
            if (!sel[c] || len[c] < s->L) continue;
            uint32_t W = len[c] - s->L + 1;
            uint32_t J = (W + PSX_STRATUM - 1) / PSX_STRATUM;
            for (uint32_t j = 0; j < J; ++j) st[q++] = (stratum_ref){so, c, j};
            uint64_t T = (uint64_t)llround(p->coverage * (double)len[c] / (double)(s->pe ? 2 * s->L : s->L));
            if (!has_ev[c]) reads += T * (s->pe ? 2u : 1u);
        }
    /* rows: segments in (output sample, chromosome) order; a segment without copy-number events holds its
     * templates' reads in template order, one with events its accepted templates' reads in template order,
     * stratum after stratum (a first pass counts each stratum's accepted templates) */
    uint64_t* st_row = (uint64_t*)calloc(n_strata ? n_strata : 1, sizeof(uint64_t));
    uint64_t row_total = 0;
    if (rows) {
        uint64_t* acc = (uint64_t*)calloc(n_strata ? n_strata : 1, sizeof(uint64_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8)
#endif
        for (int64_t qi = 0; qi < (int64_t)n_strata; ++qi) {
            uint32_t so = st[qi].so, c = st[qi].c, j = st[qi].j;
            if (!has_ev[c]) continue;
            int is_normal = so == n_req;
            const uint32_t* sl = is_normal ? NULL : f->sample_leaves + f->sample_begin[req[so]];
            uint32_t ns = is_normal ? 0 : f->sample_begin[req[so] + 1] - f->sample_begin[req[so]];
            uint32_t W = len[c] - s->L + 1, K = kmax[c];
            uint64_t T = (uint64_t)llround(p->coverage * (double)len[c] / (double)(s->pe ? 2 * s->L : s->L) *
                                           ((double)K / (double)ploidy));
            uint32_t b = j * PSX_STRATUM, e = b + PSX_STRATUM < W ? b + PSX_STRATUM : W;
            uint64_t nj = T * e / W - T * b / W;
            uint32_t c2 = (c << 16) | so;
            for (uint32_t r = 0; r < (uint32_t)nj; ++r) {
                uint32_t w[4], leaf, allele;
                philox4(seed, r, j, c2, (uint32_t)P_TEMPLATE << 24, w);
                uint32_t x = b + mulhi32(w[0], e - b);
                acc[qi] += (uint64_t)template_source(s, c, 1, K, is_normal, thr_normal, sl, ns, w, x, &leaf, &allele);
            }
        }
        uint32_t rpt = s->pe ? 2u : 1u;
        for (uint64_t qi = 0; qi < n_strata; ++qi) {
            uint32_t so = st[qi].so, c = st[qi].c, j = st[qi].j;
            if (j == 0) {  /* a segment's first stratum */
                st_row[qi] = row_total;
                if (!has_ev[c]) {
                    uint64_t T = (uint64_t)llround(p->coverage * (double)len[c] / (double)(s->pe ? 2 * s->L : s->L));
                    row_total += T * rpt;
                }
                (void)so;
            } else {
                st_row[qi] = has_ev[c] ? row_total : st_row[qi - 1];
            }
            if (has_ev[c]) row_total += acc[qi] * rpt;
        }
        free(acc);
        if (row_total > capacity) {
            free(st); free(st_row); free(req); free(s);
            return fail("the read buffer holds %llu reads, %llu needed", (unsigned long long)capacity,
                        (unsigned long long)row_total);
        }
    }
    uint64_t accepted = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : accepted)
#endif
    for (int64_t qi = 0; qi < (int64_t)n_strata; ++qi) {
        uint32_t so = st[qi].so, c = st[qi].c, j = st[qi].j;
        int is_normal = so == n_req;
        const uint32_t* sl = is_normal ? NULL : f->sample_leaves + f->sample_begin[req[so]];
        uint32_t ns = is_normal ? 0 : f->sample_begin[req[so] + 1] - f->sample_begin[req[so]];
        uint32_t W = len[c] - s->L + 1;
        /* with copy-number events, one of kmax copy slots per cell is drawn,
         * kmax / ploidy times as many templates */
        uint32_t K = kmax[c];
        uint64_t T = (uint64_t)llround(p->coverage * (double)len[c] / (double)(s->pe ? 2 * s->L : s->L) *
                                       ((double)K / (double)ploidy));
        uint32_t b = j * PSX_STRATUM, e = b + PSX_STRATUM < W ? b + PSX_STRATUM : W;
        uint64_t nj = T * e / W - T * b / W;
        uint32_t c2 = (c << 16) | so;
        int32_t* a_out = alt + (uint64_t)so * n_mut;
        int32_t* d_out = depth + (uint64_t)so * n_mut;
        uint32_t rpt = s->pe ? 2u : 1u;
        uint32_t eb[PSX_ERR_KCAP], es[PSX_ERR_KCAP], eoff[2][PSX_ERR_KCAP], esub[2][PSX_ERR_KCAP];
        uint64_t kept = 0;  /* accepted templates of the stratum so far (copy-number events) */
        for (uint32_t r = 0; r < (uint32_t)nj; ++r) {
            uint32_t w[4];
            philox4(seed, r, j, c2, (uint32_t)P_TEMPLATE << 24, w);
            uint32_t x = b + mulhi32(w[0], e - b);
            uint32_t leaf, allele;
            if (!template_source(s, c, has_ev[c], K, is_normal, thr_normal, sl, ns, w, x, &leaf, &allele)) continue;
            if (has_ev[c]) accepted += s->pe ? 2u : 1u;
            /* the template's errors per mate, in draw order */
            uint32_t ne[2] = {0, 0};
            if (s->e > 0.0) {
                uint32_t n = template_errors(s, r, j, c2, eb, es);
                for (uint32_t q = 0; q < n; ++q) {
                    uint32_t m = eb[q] / s->L;
                    eoff[m][ne[m]] = eb[q] % s->L;
                    esub[m][ne[m]] = es[q];
                    ++ne[m];
                }
            }
            /* read index (rows) */
            uint64_t ri = st_row[qi] + (has_ev[c] ? kept : T * b / W + r) * rpt;
            ++kept;
            if (s->hook) s->hook(s->hook_user, so, c, x, leaf, allele, 0, eoff[0], esub[0], ne[0]);
            else seq_read(s, c, x, leaf, allele, eoff[0], esub[0], ne[0], a_out, d_out);
            if (rows) read_row(s, c, x, leaf, allele, 0, eoff[0], esub[0], ne[0], rows + ri * s->rw);
            if (s->pe) {
                uint32_t I = s->ins_k0 + pick(s->ins_T, s->ins_n, w[3]);
                if (I < s->L) I = s->L;
                uint64_t x2 = (uint64_t)x + I - s->L;
                if (x2 > len[c] - s->L) x2 = len[c] - s->L;
                if (s->hook) s->hook(s->hook_user, so, c, (uint32_t)x2, leaf, allele, 1, eoff[1], esub[1], ne[1]);
                else seq_read(s, c, (uint32_t)x2, leaf, allele, eoff[1], esub[1], ne[1], a_out, d_out);
                if (rows) read_row(s, c, (uint32_t)x2, leaf, allele, 1, eoff[1], esub[1], ne[1], rows + (ri + 1) * s->rw);
            }
        }
    }
    free(st); free(st_row); free(req); free(s);
    *n_reads = reads + accepted;
    return 0;
}

int orc_simulate_seq_cna(uint32_t n_chr, const uint32_t* len, const orc_forest* f,
                         const psx_mutation* m, uint64_t n_mut, uint32_t ploidy,
                         const psx_cna_event* ev, uint32_t n_ev,
                         const psx_seq_params* p, uint64_t seed, int32_t* alt, int32_t* depth,
                         uint8_t* present, uint64_t* n_reads) {
    return seq_run(n_chr, len, NULL, f, m, n_mut, ploidy, ev, n_ev, p, seed, alt, depth, present, NULL, 0, NULL, NULL,
                   n_reads);
}

int orc_simulate_reads_cna(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, const orc_forest* f,
                           const psx_mutation* m, uint64_t n_mut, uint32_t ploidy, const psx_cna_event* ev,
                           uint32_t n_ev, const psx_seq_params* p, uint64_t seed, uint32_t* rows, uint64_t capacity,
                           uint64_t* n_reads) {
    uint32_t n_out = (p->samples ? p->n_samples : f->n_samples) + (p->with_normal_sample ? 1u : 0u);
    size_t n = (size_t)n_out * (n_mut ? n_mut : 1);
    int32_t* alt = (int32_t*)malloc(n * sizeof(int32_t));
    int32_t* depth = (int32_t*)malloc(n * sizeof(int32_t));
    uint8_t* present = (uint8_t*)malloc(n);
    int rc = seq_run(n_chr, len, packed, f, m, n_mut, ploidy, ev, n_ev, p, seed, alt, depth, present, rows, capacity,
                     NULL, NULL, n_reads);
    free(alt); free(depth); free(present);
    return rc;
}

int orc_simulate_reads(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, const orc_forest* f,
                       const psx_mutation* m, uint64_t n_mut, uint32_t ploidy, const psx_seq_params* p,
                       uint64_t seed, uint32_t* rows, uint64_t capacity, uint64_t* n_reads) {
    return orc_simulate_reads_cna(n_chr, len, packed, f, m, n_mut, ploidy, NULL, 0, p, seed, rows, capacity, n_reads);
}

int orc_seq_walk(uint32_t n_chr, const uint32_t* len, const orc_forest* f, const psx_mutation* m, uint64_t n_mut,
                 uint32_t ploidy, const psx_seq_params* p, uint64_t seed, orc_read_fn fn, void* user,
                 uint64_t* n_reads) {
    uint32_t n_out = (p->samples ? p->n_samples : f->n_samples) + (p->with_normal_sample ? 1u : 0u);
    size_t n = (size_t)n_out * (n_mut ? n_mut : 1);
    int32_t* alt = (int32_t*)malloc(n * sizeof(int32_t));
    int32_t* depth = (int32_t*)malloc(n * sizeof(int32_t));
    uint8_t* present = (uint8_t*)malloc(n);
    int rc = seq_run(n_chr, len, NULL, f, m, n_mut, ploidy, NULL, 0, p, seed, alt, depth, present, NULL, 0, fn, user,
                     n_reads);
    free(alt); free(depth); free(present);
    return rc;
}

/* ------------------------------------------------------------- result table */

/* get_active_mutations + add_sample_statistics (src/seq_simulation.cpp:92-167):
 * a row per SID with an alternative read in some sample (or, when
 * include_non_sequenced, held by some sample); VAF = occ/cov when the sample
 * holds the SID, 0 otherwise (:123-130). */
int orc_seq_table(const psx_mutation* recs, uint64_t n, uint32_t S, const int32_t* alt,
                  const int32_t* depth, const uint8_t* present, int32_t include,
                  uint64_t* n_rows, uint64_t* row_rec, int32_t* occ, int32_t* cov, double* vaf) {
    uint64_t rows = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (!recs[i].sid_head) continue;
        uint64_t g1 = i + 1;
        while (g1 < n && !recs[g1].sid_head) ++g1;
        int any_occ = 0, any_pres = 0;
        for (uint32_t s = 0; s < S; ++s) {
            any_occ |= alt[(uint64_t)s * n + i] > 0;
            for (uint64_t q = i; q < g1; ++q) any_pres |= present[(uint64_t)s * n + q] != 0;
        }
        if (!(any_occ || (include && any_pres))) continue;
        row_rec[rows] = i;
        for (uint32_t s = 0; s < S; ++s) {
            int32_t o = alt[(uint64_t)s * n + i], cv = depth[(uint64_t)s * n + i];
            int in_data = o > 0;
            for (uint64_t q = i; q < g1; ++q) in_data |= present[(uint64_t)s * n + q] != 0;
            occ[(uint64_t)s * n + rows] = in_data ? o : 0;
            cov[(uint64_t)s * n + rows] = cv;
            vaf[(uint64_t)s * n + rows] = in_data ? (double)o / (double)cv : 0.0;
        }
        ++rows;
    }
    *n_rows = rows;
    return 0;
}
