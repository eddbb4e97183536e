// simulate_seq on gfx950: the ReadSimulator call of the reference
// (src/seq_simulation.cpp:354-429, RACES call at :371-372) in count-only
// mode, its read-materialising analogue (write_SAM = TRUE, :564) and the
// batched form. Kernels are in psx_seq_dev.h; this file builds descriptors
// on the host, launches, and times the dominant kernel with HIP events.
#include <algorithm>
#include <numeric>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>

#include "psx_seq_dev.h"
#include "psx_seq_host.h"

using namespace psxk;

namespace psxh {

struct seq_batch {
    std::vector<job_dev> jobs;
    std::vector<seg_dev> segs;
    std::vector<pres_dev> pres;
    uint64_t n_reads = 0, pres_total = 0;
    uint64_t n_items = 0;  // written on the device (k_seq_items) from the segments
    uint64_t n_items_sparse = 0;  // the first items: those of sparse jobs (layout)
    uint32_t spi = 1;      // strata per item
    bool no_alloc = false;  // refuse jobs whose mutation table lacks an index for their genome
    scratch_builder sb;
    std::vector<size_t> tab_off;  // per job: insert thresholds, insert guide, error thresholds
    std::vector<size_t> cna_off;  // per job: copy-number events (SIZE_MAX: none)
    std::vector<uint64_t> job_reads;  // per job: reads of the sampling model without copy-number events
    std::vector<double> job_cover;  // per job: expected covered positions per 64 templates (0: no errors)
    bool cna_sparse = false, cna_dense = false;  // copy-number events among the sparse / the other jobs
    size_t o_jobs = 0, o_segs = 0, o_items = 0, o_pres = 0, o_acc = 0;
    uint8_t* dev = nullptr;       // device base of the last upload
};

// The compact insert-size and error-count tables depend only on their
// parameters, jobs of a batch and requests of a server mostly share them, and
// building one takes tens of microseconds, so the process keeps up to 256 of
// each (copies are handed out; failures are not kept).
struct table_cache {
    std::mutex mu;
    std::map<std::pair<uint32_t, uint32_t>, psx_ctable> ins;  // (mean, sd)
    std::map<std::pair<uint32_t, uint64_t>, psx_ctable> err;  // (bases per template, bits of the rate)
};
static table_cache& tables() {
    static table_cache c;
    return c;
}

static psx_ctable insert_ctable(uint32_t mean, uint32_t sd) {
    table_cache& c = tables();
    const auto key = std::make_pair(mean, sd);
    {
        std::lock_guard<std::mutex> g(c.mu);
        auto it = c.ins.find(key);
        if (it != c.ins.end()) return it->second;
    }
    psx_ctable t = psx_compact_table(psx_insert_table(mean, sd));  // throws on invalid parameters
    std::lock_guard<std::mutex> g(c.mu);
    if (c.ins.size() >= 256) c.ins.clear();
    c.ins.emplace(key, t);
    return t;
}

// The error-count table of N bases per template at rate e; throws when the
// rate is too high for N.
static psx_ctable error_ctable(uint32_t N, double e, uint32_t L) {
    table_cache& c = tables();
    uint64_t bits;
    std::memcpy(&bits, &e, sizeof bits);
    const auto key = std::make_pair(N, bits);
    {
        std::lock_guard<std::mutex> g(c.mu);
        auto it = c.err.find(key);
        if (it != c.err.end()) return it->second;
    }
    if (!psx_error_rate_ok(N, e)) psx_throw(PSX_EINVAL, "error rate too high for %u-base reads", L);
    psx_ctable t = psx_compact_table(psx_error_table(N, e));
    std::lock_guard<std::mutex> g(c.mu);
    if (c.err.size() >= 256) c.err.clear();
    c.err.emplace(key, t);
    return t;
}

// Largest copy number of any leaf on a chromosome (at least the ploidy): the
// copy numbers change only at event boundaries, in leaf rank and position.
static uint32_t cna_kmax(const std::vector<psx_cna_dev>& ev, uint32_t n_leaves, uint32_t len, uint32_t ploidy) {
    std::vector<uint32_t> leaves{0}, xs{0};
    for (const psx_cna_dev& e : ev) {
        leaves.push_back(e.lo);
        if (e.hi < n_leaves) leaves.push_back(e.hi);
        xs.push_back(e.start);
        if (e.end < len) xs.push_back(e.end);
    }
    uint32_t best = ploidy, cn[PSX_MAX_CNA_PLOIDY];
    for (uint32_t l : leaves) {
        if (l >= n_leaves) continue;
        for (uint32_t x : xs) best = std::max(best, psx_cna_copies(ev.data(), (uint32_t)ev.size(), l, x, ploidy, cn));
    }
    return best;
}

bool batch_has_cna(const seq_batch* b);

uint32_t out_samples(const psx_forest* f, const psx_seq_params* p) {
    return (p->samples ? p->n_samples : f->n_samples) + (p->with_normal_sample ? 1u : 0u);
}

namespace {

// Validates one job and appends its descriptors; the ordering of output
// samples and chromosomes follows the oracle. Nothing is appended when the
// job is invalid.
void prepare_job(psx_ctx* ctx, seq_batch& P, const job_in& in, uint32_t spi) {
    const psx_seq_params* p = in.p;
    const psx_forest* f = in.f;
    psx_mutations* m = const_cast<psx_mutations*>(in.m);
    const psx_genome* g = in.g;
    if (!(p->coverage >= 0.0) || !std::isfinite(p->coverage)) psx_throw(PSX_EINVAL, "invalid coverage");
    if (p->read_size < 1 || p->read_size > MAX_READ)
        psx_throw(PSX_EINVAL, "the read size must be in [1, %u]", MAX_READ);
    if (!(p->purity >= 0.0 && p->purity <= 1.0)) psx_throw(PSX_EINVAL, "The purity must be in [0,1]");
    if (!(p->error_rate >= 0.0 && p->error_rate < 1.0))
        psx_throw(PSX_EINVAL, "The parameter \"error_rate\" must be a positive real number.");
    if (m->n_leaves != f->n_leaves) psx_throw(PSX_EINVAL, "the mutation table belongs to another forest");
    const uint32_t ploidy = m->ploidy;
    if ((uint64_t)f->n_leaves * ploidy >= (uint64_t(1) << 32)) psx_throw(PSX_EINVAL, "too many cell copies");
    psx_cna_check(m->cna, f->n_leaves, ploidy, g->len);
    const uint32_t L = p->read_size;
    const bool pe = p->insert_size_mean > 0;
    psx_ctable ins, err;
    uint32_t reach = L;
    if (pe) {
        ins = insert_ctable(p->insert_size_mean, p->insert_size_stddev);
        if (ins.T.size() > MAX_INS)
            psx_throw(PSX_EINVAL, "the insert-size distribution is too wide (sd %u)", p->insert_size_stddev);
        reach = std::max(reach, ins.k0 + (uint32_t)ins.T.size() - 1);
        if (reach > MAX_READ) psx_throw(PSX_EINVAL, "insert sizes above %u are not supported", MAX_READ);
    }
    if (p->error_rate > 0.0) err = error_ctable((pe ? 2u : 1u) * L, p->error_rate, L);
    std::vector<uint8_t> sel(g->n_chr, p->n_chromosomes == 0 ? 1 : 0);
    for (uint32_t k = 0; k < p->n_chromosomes; ++k) {
        if (p->chromosomes[k] >= g->n_chr) psx_throw(PSX_EINVAL, "unknown chromosome %u", p->chromosomes[k]);
        sel[p->chromosomes[k]] = 1;
    }
    const uint32_t n_req = p->samples ? p->n_samples : f->n_samples;
    std::vector<uint32_t> req(n_req);
    for (uint32_t k = 0; k < n_req; ++k) {
        req[k] = p->samples ? p->samples[k] : k;
        if (req[k] >= f->n_samples) psx_throw(PSX_EINVAL, "unknown sample %u", req[k]);
        if (f->sample_begin[req[k] + 1] == f->sample_begin[req[k]])
            psx_throw(PSX_EINVAL, "sample %u has no cells", req[k]);
    }
    const psx_mut_index* ix = P.no_alloc ? psx_mutations_find_index(m, g->len) : &psx_mutations_index(ctx, m, g->len);
    if (!ix)
        psx_throw(PSX_EINVAL, "the mutation table has no index for this genome's chromosome lengths yet (an "
                              "allocating sequencing call builds it)");
    const uint32_t n_out = n_req + (p->with_normal_sample ? 1u : 0u);
    const uint32_t job = (uint32_t)P.jobs.size();
    job_dev J{};
    J.seed = in.seed;
    J.pos = m->pos; J.lo = m->lo; J.hi = m->hi;
    J.info = m->info; J.bucket = ix->bucket;
    J.sample_leaves = f->d_sample_leaves.as<uint32_t>();
    J.leaf_count = f->d_leaf_count.as<uint4>();
    J.leaf_blocks = f->leaf_blocks;
    J.alt = in.alt; J.depth = in.depth; J.present = in.present;
    J.packed = g->packed.as<uint32_t>();
    J.n_mut = m->n;
    J.thr_normal = (uint64_t)((1.0 - p->purity) * 4294967296.0);
    J.ins_k0 = ins.k0; J.ins_n = (uint32_t)ins.T.size();
    J.err_n = (uint32_t)err.T.size();
    J.err_k0 = err.k0;
    J.err_N = (pe ? 2u : 1u) * L;
    J.inv_L = (uint32_t)(((uint64_t(1) << 32) + L - 1) / L);
    J.L = L; J.pe = pe ? 1u : 0u; J.ploidy = ploidy; J.n_leaves = f->n_leaves;
    J.preneo = p->preneoplastic_in_normal ? 1u : 0u; J.reach = reach; J.n_out = n_out;
    J.purity_lt1 = p->purity < 1.0 ? 1u : 0u;
    P.tab_off.push_back(ins.T.empty() ? SIZE_MAX : P.sb.put_vec(ins.T));
    P.tab_off.push_back(ins.guide.empty() ? SIZE_MAX : P.sb.put_vec(ins.guide));
    P.tab_off.push_back(err.T.empty() ? SIZE_MAX : P.sb.put_vec(err.T));
    // copy-number events per chromosome (doublings on every chromosome), in DFS order
    std::vector<psx_cna_dev> cdev;
    std::vector<uint32_t> c_off(g->n_chr, 0), c_n(g->n_chr, 0), c_kmax(g->n_chr, ploidy);
    for (uint32_t c = 0; c < g->n_chr && !m->cna.empty(); ++c) {
        std::vector<psx_cna_dev> ev;
        for (const psx_cna_event& e : m->cna)
            if (e.chr == c || e.chr == PSX_ALL_CHROMOSOMES)
                ev.push_back(psx_cna_dev{e.leaf_lo, e.leaf_hi, e.chr == c ? e.start : 0u,
                                         e.chr == c ? e.end : g->len[c], psx_cna_code(e)});
        c_off[c] = (uint32_t)cdev.size();
        c_n[c] = (uint32_t)ev.size();
        c_kmax[c] = cna_kmax(ev, f->n_leaves, g->len[c], ploidy);
        if ((uint64_t)f->n_leaves * c_kmax[c] >= (uint64_t(1) << 32))
            psx_throw(PSX_EINVAL, "too many cell copies (%u copies on chromosome %u)", c_kmax[c], c);
        if (c_kmax[c] > 0xFFFFu)
            psx_throw(PSX_EINVAL, "copy numbers above 65535 are not supported (%u on chromosome %u)", c_kmax[c], c);
        cdev.insert(cdev.end(), ev.begin(), ev.end());
    }
    P.cna_off.push_back(cdev.empty() ? SIZE_MAX : P.sb.put_vec(cdev));
    P.jobs.push_back(J);
    double cover = 0.0;
    if (p->error_rate > 0.0) {
        // expected positions covered by 64 templates: records per base x bases read, times the share of
        // templates kept under copy numbers (ploidy / kmax, length-weighted)
        double recs = 0.0, bases = 0.0, kept = 0.0;
        for (uint32_t c = 0; c < g->n_chr; ++c) {
            if (!sel[c] || g->len[c] < L) continue;
            recs += (double)(m->chr_begin[c + 1] - m->chr_begin[c]);
            bases += (double)g->len[c];
            kept += (double)g->len[c] * (double)ploidy / (double)c_kmax[c];
        }
        if (bases > 0.0) cover = 64.0 * (pe ? 2.0 : 1.0) * L * recs / bases * kept / bases;
    }
    P.job_cover.push_back(cover);
    const uint64_t reads_before = P.n_reads;
    for (uint32_t so = 0; so < n_out; ++so) {
        const bool normal = so == n_req;
        const uint32_t leaf_off = normal ? 0u : f->sample_begin[req[so]];
        const uint32_t n_cells = normal ? 0u : f->sample_begin[req[so] + 1] - f->sample_begin[req[so]];
        if (m->n) {
            P.pres.push_back(pres_dev{P.pres_total, job, so, leaf_off, n_cells, normal ? 0u : req[so], 0u});
            P.pres_total += m->n;
        }
        for (uint32_t c = 0; c < g->n_chr; ++c) {
            if (!sel[c] || g->len[c] < L) continue;
            seg_dev S{};
            S.n_cna = c_n[c];
            S.cna_off = c_off[c];
            S.kmax = c_kmax[c];
            // with copy-number events a template draws one of kmax copy slots per cell
            const double t_real = p->coverage * (double)g->len[c] / (double)(pe ? 2 * L : L) *
                                  ((double)S.kmax / (double)ploidy);
            if (!(t_real < 9.0e18)) psx_throw(PSX_EINVAL, "too many templates on chromosome %u", c);
            S.T = (uint64_t)std::llround(t_real);
            S.out_off = (uint64_t)so * m->n;
            S.read_base = P.n_reads;
            S.chr_woff = g->woff[c];
            S.job = job; S.so = so; S.chr = c; S.len = g->len[c];
            S.W = g->len[c] - L + 1;
            S.J = (S.W + PSX_STRATUM - 1) / PSX_STRATUM;
            if (S.T && S.W > UINT64_MAX / S.T)  // stratum bounds are T * start / W in 64 bits
                psx_throw(PSX_EINVAL, "too many templates on chromosome %u (%llu over %u starts)", c,
                          (unsigned long long)S.T, S.W);
            S.bucket_off = (uint32_t)ix->bucket_off[c];
            S.n_bucket = (uint32_t)(ix->bucket_off[c + 1] - ix->bucket_off[c]);
            S.leaf_off = leaf_off; S.n_cells = n_cells;
            P.n_reads += S.T * (pe ? 2u : 1u);  // template draws: an upper bound with copy-number events
            P.segs.push_back(S);  // its work items are numbered in layout()
        }
    }
    P.job_reads.push_back(P.n_reads - reads_before);
}

// A job whose reads cover fewer than 48 positions per 64 templates (or that
// draws no errors) counts with errors checked per covered position.
bool job_sparse(const seq_batch& P, uint32_t job) { return P.job_cover[job] < 48.0; }

// Numbers the work items (the sparse jobs' segments first, so that each kind
// of job is one range of items for its own kernel) and places the descriptors
// of a prepared batch after the per-job tables.
void layout(seq_batch& P) {
    std::stable_partition(P.segs.begin(), P.segs.end(), [&](const seg_dev& S) { return job_sparse(P, S.job); });
    P.n_items = 0;
    for (seg_dev& S : P.segs) {
        S.first_item = (uint32_t)P.n_items;
        if (S.T) P.n_items += (S.J + P.spi - 1) / P.spi;
        if (P.n_items >= (uint64_t(1) << 32)) psx_throw(PSX_EINVAL, "too many work items in one batch");
        if (job_sparse(P, S.job)) P.n_items_sparse = P.n_items;
    }
    for (uint32_t k = 0; k < (uint32_t)P.jobs.size(); ++k)
        if (P.cna_off[k] != SIZE_MAX) (job_sparse(P, k) ? P.cna_sparse : P.cna_dense) = true;
    const std::vector<uint64_t> zeros(P.jobs.size(), 0);
    P.o_acc = P.sb.put_vec(zeros);
    P.o_segs = P.sb.put_vec(P.segs);
    P.o_pres = P.sb.put_vec(P.pres);
    P.o_jobs = P.sb.put_vec(P.jobs);
    P.o_items = (P.sb.host.size() + 255) & ~size_t(255);  // device-only, after the uploaded descriptors
}

template <int BLOCK, int CAP>
void launch_count_t(const seq_lane& lane, const seq_launch& L, uint32_t grid, bool cna, bool defer) {
    if (cna && defer)
        hipLaunchKernelGGL((k_seq_count<BLOCK, CAP, true, true>), dim3(grid), dim3(BLOCK), 0, lane.stream, L);
    else if (cna)
        hipLaunchKernelGGL((k_seq_count<BLOCK, CAP, true, false>), dim3(grid), dim3(BLOCK), 0, lane.stream, L);
    else if (defer)
        hipLaunchKernelGGL((k_seq_count<BLOCK, CAP, false, true>), dim3(grid), dim3(BLOCK), 0, lane.stream, L);
    else
        hipLaunchKernelGGL((k_seq_count<BLOCK, CAP, false, false>), dim3(grid), dim3(BLOCK), 0, lane.stream, L);
}

void launch_count(const seq_lane& lane, const psx_launch_config& c, const seq_launch& L, uint32_t grid, bool cna,
                  bool defer) {
#define PSX_CASE(B, C) \
    if (c.block_threads == B && c.lds_records == C) return launch_count_t<B, C>(lane, L, grid, cna, defer);
    PSX_CASE(256, 512) PSX_CASE(256, 1024) PSX_CASE(256, 2048)
    PSX_CASE(512, 512) PSX_CASE(512, 1024) PSX_CASE(512, 2048)
#undef PSX_CASE
    psx_throw(PSX_EINVAL, "unsupported launch configuration");
}

seq_launch launch_desc(const seq_batch& P, uint8_t* base, unsigned long long* queue) {
    seq_launch L{};
    L.jobs = (const job_dev*)(base + P.o_jobs);
    L.segs = (const seg_dev*)(base + P.o_segs);
    L.items = (const item_dev*)(base + P.o_items);
    L.n_items = P.n_items;
    L.queue = queue;
    return L;
}

}  // namespace

seq_batch* batch_prepare(psx_ctx* ctx, const job_in* jobs, size_t n_jobs, uint32_t spi, bool no_alloc) {
    auto P = std::make_unique<seq_batch>();
    P->spi = spi;
    P->no_alloc = no_alloc;
    for (size_t k = 0; k < n_jobs; ++k) prepare_job(ctx, *P, jobs[k], spi);
    layout(*P);
    return P.release();
}

void batch_free(seq_batch* b) { delete b; }
size_t batch_desc_bytes(const seq_batch* b) { return b->o_items + b->n_items * sizeof(item_dev) + 256; }
uint64_t batch_reads(const seq_batch* b) { return b->n_reads; }
bool batch_has_cna(const seq_batch* b) {
    for (size_t off : b->cna_off)
        if (off != SIZE_MAX) return true;
    return false;
}

std::vector<uint64_t> batch_job_reads(const seq_batch* b, const seq_lane& lane) {
    std::vector<uint64_t> r = b->job_reads;
    if (!batch_has_cna(b)) return r;
    std::vector<uint64_t> acc(b->jobs.size());
    PSX_HIP(hipMemcpyAsync(acc.data(), lane.desc + b->o_acc, 8 * acc.size(), hipMemcpyDeviceToHost, lane.stream));
    PSX_HIP(hipStreamSynchronize(lane.stream));
    for (size_t k = 0; k < r.size(); ++k)
        if (b->cna_off[k] != SIZE_MAX) r[k] = acc[k];
    return r;
}
size_t batch_items(const seq_batch* b) { return b->n_items; }

void batch_upload(seq_batch* P, const seq_lane& lane) {
    if (lane.desc_bytes < batch_desc_bytes(P)) psx_throw(PSX_EINVAL, "descriptor buffer too small");
    uint8_t* base = lane.desc;
    job_dev* hj = (job_dev*)(P->sb.host.data() + P->o_jobs);
    auto at = [&](size_t off) -> const void* { return off == SIZE_MAX ? nullptr : base + off; };
    for (size_t k = 0; k < P->jobs.size(); ++k) {
        hj[k].ins_T = (const uint32_t*)at(P->tab_off[3 * k]);
        hj[k].ins_guide = (const uint16_t*)at(P->tab_off[3 * k + 1]);
        hj[k].err_T = (const uint32_t*)at(P->tab_off[3 * k + 2]);
        hj[k].cna = (const psx_cna_dev*)at(P->cna_off[k]);
        hj[k].acc = P->cna_off[k] == SIZE_MAX ? nullptr : (unsigned long long*)(base + P->o_acc) + k;
    }
    if (!P->sb.host.empty())
        PSX_HIP(hipMemcpyAsync(base, P->sb.host.data(), P->sb.host.size(), hipMemcpyHostToDevice, lane.stream));
    if (P->n_items) {
        const uint32_t grid = (uint32_t)((P->n_items + 255) / 256);
        hipLaunchKernelGGL(k_seq_items, dim3(grid), dim3(256), 0, lane.stream, (const seg_dev*)(base + P->o_segs),
                           (uint32_t)P->segs.size(), (item_dev*)(base + P->o_items), (uint32_t)P->n_items, P->spi);
        PSX_HIP(hipGetLastError());
    }
    P->dev = base;
}

void batch_launch(const seq_batch* P, const seq_lane& lane, const psx_launch_config& cfg, bool timed) {
    if (P->dev != lane.desc) psx_throw(PSX_EINVAL, "descriptors were uploaded to another lane");
    const seq_launch L = launch_desc(*P, lane.desc, lane.queue);
    const pres_dev* ps = (const pres_dev*)(lane.desc + P->o_pres);
    const uint64_t want = (uint64_t)lane.n_cu * cfg.blocks_per_cu;
    const uint32_t init_grid = (uint32_t)std::max<uint64_t>(1, (P->pres_total + 255) / 256);
    hipLaunchKernelGGL(k_seq_init, dim3(init_grid), dim3(256), 0, lane.stream, L.jobs, ps, (uint32_t)P->pres.size(),
                       P->pres_total, L.queue, (uint32_t)P->jobs.size());
    if (timed) PSX_HIP(hipEventRecord(lane.ev0, lane.stream));
    // the sparse jobs' items (errors checked per covered position), then the others' (errors per template), each
    // range from its own counter
    const uint64_t ns = P->n_items_sparse, nd = P->n_items - ns;
    if (ns || !nd) {
        seq_launch Ls = L;
        Ls.n_items = ns;
        launch_count(lane, cfg, Ls, (uint32_t)std::max<uint64_t>(1, std::min(want, ns)), P->cna_sparse, true);
    }
    if (nd) {
        seq_launch Ld = L;
        Ld.items = L.items + ns;
        Ld.n_items = nd;
        Ld.queue = L.queue + 1;
        launch_count(lane, cfg, Ld, (uint32_t)std::max<uint64_t>(1, std::min(want, nd)), P->cna_dense, false);
    }
    PSX_HIP(hipGetLastError());
    if (timed) PSX_HIP(hipEventRecord(lane.ev1, lane.stream));
}

psx_launch_config default_config() { return psx_launch_config{256, 512, 1, 8, 0}; }

void check_config(const psx_launch_config& c) {
    const bool block_ok = c.block_threads == 256 || c.block_threads == 512;
    const bool cap_ok = c.lds_records == 512 || c.lds_records == 1024 || c.lds_records == 2048;
    if (!block_ok || !cap_ok || c.strata_per_item < 1 || c.blocks_per_cu < 1 || c.blocks_per_cu > 64)
        psx_throw(PSX_EINVAL, "unsupported launch configuration (block %u, lds %u, strata %u, blocks/CU %u)",
                  c.block_threads, c.lds_records, c.strata_per_item, c.blocks_per_cu);
}

run_result run_prepared(psx_ctx* ctx, seq_batch* batch, const seq_lane& lane, const psx_launch_config& cfg,
                        const std::function<void()>& during) {
    std::unique_ptr<seq_batch, void (*)(seq_batch*)> P(batch, batch_free);
    batch_upload(P.get(), lane);
    if (cfg.use_graph) {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        PSX_HIP(hipStreamBeginCapture(lane.stream, hipStreamCaptureModeThreadLocal));
        batch_launch(P.get(), lane, cfg, false);
        PSX_HIP(hipStreamEndCapture(lane.stream, &graph));
        PSX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        PSX_HIP(hipEventRecord(lane.ev0, lane.stream));
        PSX_HIP(hipGraphLaunch(exec, lane.stream));
        PSX_HIP(hipEventRecord(lane.ev1, lane.stream));
        if (during) during();
        PSX_HIP(hipStreamSynchronize(lane.stream));
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
    } else {
        batch_launch(P.get(), lane, cfg, true);
        if (during) during();
        PSX_HIP(hipStreamSynchronize(lane.stream));
    }
    float ms = 0.f;
    PSX_HIP(hipEventElapsedTime(&ms, lane.ev0, lane.ev1));
    ctx->last_kernel_ms = ms;
    uint64_t n = 0;
    for (uint64_t r : batch_job_reads(P.get(), lane)) n += r;
    return run_result{n, (double)ms};
}

}  // namespace psxh

using namespace psxh;

namespace {

seq_lane ctx_lane(psx_ctx* ctx, size_t desc_bytes) {
    ctx->scratch.reserve(desc_bytes);
    ctx->queue.reserve(256);
    seq_lane lane;
    lane.n_cu = ctx->n_cu;
    lane.stream = ctx->stream;
    lane.ev0 = ctx->ev0;
    lane.ev1 = ctx->ev1;
    lane.desc = (uint8_t*)ctx->scratch.p;
    lane.desc_bytes = ctx->scratch.bytes;
    lane.queue = ctx->queue.as<unsigned long long>();
    return lane;
}

// The jobs in one launch with descriptors in the context's scratch buffer.
run_result run_count(psx_ctx* ctx, const std::vector<job_in>& jobs, const psx_launch_config& cfg,
                     const std::function<void()>& during = {}) {
    check_config(cfg);
    PSX_HIP(hipSetDevice(ctx->device));
    seq_batch* P = batch_prepare(ctx, jobs.data(), jobs.size(), cfg.strata_per_item);
    seq_lane lane;
    try {
        lane = ctx_lane(ctx, batch_desc_bytes(P));
    } catch (...) {
        batch_free(P);
        throw;
    }
    return run_prepared(ctx, P, lane, cfg, during);
}

// Writes one byte per 4 KiB page of a host buffer, so its pages are mapped
// before a device copy fills it (fresh pageable memory otherwise faults page
// by page inside the copy).
void touch_pages(void* p, size_t bytes) {
    volatile uint8_t* b = (volatile uint8_t*)p;
    for (size_t i = 0; i < bytes; i += 4096) b[i] = 0;
    if (bytes) b[bytes - 1] = 0;
}

std::vector<job_in> jobs_of(const psx_seq_job* jobs, uint32_t n_jobs) {
    std::vector<job_in> in;
    for (uint32_t k = 0; k < n_jobs; ++k) {
        const psx_seq_job& J = jobs[k];
        if (!J.genome || !J.forest || !J.mutations || !J.d_alt || !J.d_depth || !J.d_present)
            psx_throw(PSX_EINVAL, "job %u is incomplete", k);
        in.push_back(job_in{J.genome, J.forest, J.mutations, &J.params, J.seed, J.d_alt, J.d_depth, J.d_present});
    }
    return in;
}

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
            total += (uint64_t)std::llround(p->coverage * (double)g->len[c] / (double)(pe ? 2 * L : L)) *
                     (pe ? 2u : 1u);
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
        psx_launch_config c = ctx->tuned ? ctx->tuned_cfg : default_config();
        c.use_graph = 0;
        run_result r = run_count(ctx, {job_in{g, f, m, p, seed, d_alt, d_depth, d_present}}, c);
        if (n_reads) *n_reads = r.n_reads;
    });
}

// fp32 VAF per (output sample, record): the SID's alternative reads over its
// depth at a SID head when the sample holds the SID, else 0 (psx_seq_table's
// rule, without the row filter).
__global__ void __launch_bounds__(256) k_seq_vaf(const uint32_t* __restrict__ info, uint64_t n, uint64_t total,
                                                 const int32_t* __restrict__ alt, const int32_t* __restrict__ depth,
                                                 const uint8_t* __restrict__ present, float* __restrict__ vaf) {
    const uint64_t q = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (q >= total) return;
    const uint64_t m = q % n, base = q - m;
    float v = 0.0f;
    if (PSX_INFO_HEAD(info[m])) {
        bool held = alt[q] > 0;
        for (uint64_t k = m; k < n && !held && (k == m || !PSX_INFO_HEAD(info[k])); ++k) held = present[base + k] != 0;
        if (held) v = (float)alt[q] / (float)depth[q];
    }
    vaf[q] = v;
}

extern "C" psx_status psx_seq_vaf_f32(psx_ctx* ctx, const psx_mutations* m, uint32_t n_out_samples,
                                      const int32_t* d_alt, const int32_t* d_depth, const uint8_t* d_present,
                                      float* d_vaf) {
    if (!ctx || !m || (m->n && (!d_alt || !d_depth || !d_present || !d_vaf))) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        const uint64_t total = (uint64_t)n_out_samples * m->n;
        if (!total) return;
        PSX_HIP(hipSetDevice(ctx->device));
        hipLaunchKernelGGL(k_seq_vaf, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, ctx->stream,
                           m->info, m->n, total, d_alt, d_depth, d_present, d_vaf);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipStreamSynchronize(ctx->stream));
    });
}

extern "C" psx_status psx_simulate_seq(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                       const psx_mutations* m, const psx_seq_params* p, uint64_t seed, int32_t* alt,
                                       int32_t* depth, uint8_t* present, uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        const uint64_t n = (uint64_t)out_samples(f, p) * m->n;
        // the device tables in the context's output buffer (grown, kept for later calls): alt | depth | present
        const size_t o_depth = (n * 4 + 255) / 256 * 256, o_present = o_depth + (n * 4 + 255) / 256 * 256;
        ctx->seq_out.reserve(o_present + n);
        uint8_t* d = (uint8_t*)ctx->seq_out.p;
        psx_launch_config c = ctx->tuned ? ctx->tuned_cfg : default_config();
        c.use_graph = 0;
        // while the kernels run, the host maps the output pages the copies will fill
        auto prefault = [&] {
            if (alt) touch_pages(alt, n * 4);
            if (depth) touch_pages(depth, n * 4);
            if (present) touch_pages(present, n);
        };
        run_result r = run_count(ctx, {job_in{g, f, m, p, seed, (int32_t*)d, (int32_t*)(d + o_depth),
                                              d + o_present}}, c, prefault);
        std::vector<psx_copy_range> out;
        if (alt) out.push_back({d, alt, n * 4});
        if (depth) out.push_back({d + o_depth, depth, n * 4});
        if (present) out.push_back({d + o_present, present, n});
        psx_copy_out(ctx, out);
        if (n_reads) *n_reads = r.n_reads;
    });
}

// One persistent read-materialising launch: as many blocks as fit on every CU
// at once (capped by the work items).
template <int BLOCK, int CAP, bool STAGE, int NW, bool CNA = false>
void launch_reads(psx_ctx* ctx, const seq_launch& L, size_t items, size_t stage) {
    int per_cu = 0;
    PSX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_seq_reads<BLOCK, CAP, STAGE, NW, CNA>, BLOCK,
                                                         stage));
    const uint64_t want = (uint64_t)ctx->n_cu * (uint64_t)std::max(per_cu, 1);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, items));
    PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL((k_seq_reads<BLOCK, CAP, STAGE, NW, CNA>), dim3(grid), dim3(BLOCK), stage, ctx->stream, L);
}

extern "C" uint32_t psx_read_words(const psx_seq_params* p) {
    return p ? 2u + (p->read_size + 15u) / 16u : 0u;
}

namespace {

// Reads of a prepared batch in output order when some segments have
// copy-number events: there only the kept templates have reads, so their
// count comes from the device. Before the upload (prepare) every segment with
// events gets the index of its first stratum and every job the device array
// of those strata's first reads; after it (finish) the kept templates of each
// stratum are counted (k_cn_accept), every segment's first read and each such
// stratum's first read are set in output order, and the segments are
// uploaded again. Without events nothing changes.
struct cn_rows {
    dbuf buf;                    // kept | first read (uint64 per stratum) | segment list | first strata
    std::vector<uint32_t> seg, first;
    uint64_t n = 0;
    std::vector<uint64_t> seg_reads;  // reads of each segment (finish)
    bool prepare(seq_batch& P) {
        seg_dev* hs = (seg_dev*)(P.sb.host.data() + P.o_segs);
        first.assign(1, 0);
        for (uint32_t k = 0; k < (uint32_t)P.segs.size(); ++k)
            if (hs[k].n_cna && hs[k].T) {
                hs[k].st_off = (uint32_t)n;
                seg.push_back(k);
                n += hs[k].J;
                first.push_back((uint32_t)n);
            }
        if (!n) return false;
        if (n >= (uint64_t(1) << 32)) psx_throw(PSX_EINVAL, "too many strata with copy-number events");
        buf.alloc(16 * n + 4 * (seg.size() + first.size()) + 256);
        job_dev* hj = (job_dev*)(P.sb.host.data() + P.o_jobs);
        for (size_t k = 0; k < P.jobs.size(); ++k) hj[k].cn_row = buf.as<uint64_t>() + n;
        return true;
    }
    void finish(psx_ctx* ctx, seq_batch& P, const seq_lane& lane) {
        uint64_t* d_kept = buf.as<uint64_t>();
        uint32_t* d_seg = (uint32_t*)(d_kept + 2 * n);
        uint32_t* d_first = d_seg + seg.size();
        PSX_HIP(hipMemcpyAsync(d_seg, seg.data(), 4 * seg.size(), hipMemcpyHostToDevice, lane.stream));
        PSX_HIP(hipMemcpyAsync(d_first, first.data(), 4 * first.size(), hipMemcpyHostToDevice, lane.stream));
        psxk::accept_args a{(const job_dev*)(lane.desc + P.o_jobs), (const seg_dev*)(lane.desc + P.o_segs), d_seg,
                            d_first, (uint32_t)seg.size(), d_kept};
        hipLaunchKernelGGL(psxk::k_cn_accept, dim3((uint32_t)((n + 3) / 4)), dim3(256), 0, lane.stream, a);
        PSX_HIP(hipGetLastError());
        std::vector<uint64_t> kept(n), row(n);
        PSX_HIP(hipMemcpyAsync(kept.data(), d_kept, 8 * n, hipMemcpyDeviceToHost, lane.stream));
        PSX_HIP(hipStreamSynchronize(lane.stream));
        // output order = the order of the first reads the segments got at preparation (job, sample, chromosome)
        seg_dev* hs = (seg_dev*)(P.sb.host.data() + P.o_segs);
        const job_dev* hj = (const job_dev*)(P.sb.host.data() + P.o_jobs);
        std::vector<uint32_t> order(P.segs.size());
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a2, uint32_t b2) { return hs[a2].read_base < hs[b2].read_base; });
        std::fill(P.job_reads.begin(), P.job_reads.end(), 0);
        seg_reads.assign(P.segs.size(), 0);
        uint64_t at = 0;
        for (uint32_t k : order) {
            seg_dev& S = hs[k];
            const uint64_t rpt = hj[S.job].pe ? 2u : 1u;
            S.read_base = at;
            uint64_t reads = S.T * rpt;
            if (S.n_cna && S.T) {
                uint64_t acc = 0;
                for (uint32_t j = 0; j < S.J; ++j) {
                    row[S.st_off + j] = at + acc * rpt;
                    acc += kept[S.st_off + j];
                }
                reads = acc * rpt;
            }
            at += reads;
            seg_reads[k] = reads;
            P.job_reads[S.job] += reads;
        }
        P.n_reads = at;
        PSX_HIP(hipMemcpyAsync(d_kept + n, row.data(), 8 * n, hipMemcpyHostToDevice, lane.stream));
        PSX_HIP(hipMemcpyAsync(lane.desc + P.o_segs, hs, sizeof(seg_dev) * P.segs.size(), hipMemcpyHostToDevice,
                               lane.stream));
        PSX_HIP(hipStreamSynchronize(lane.stream));  // the host arrays above go out of scope
    }
};

// Read materialisation of jobs in one persistent launch: job k's reads start
// at read index sum of the earlier jobs' reads (segments keep their job-order
// read_base whatever order their work items take).
void run_reads(psx_ctx* ctx, const std::vector<job_in>& in, uint32_t* d_reads, uint64_t capacity_reads,
               uint64_t* job_first, uint64_t* n_reads) {
    PSX_HIP(hipSetDevice(ctx->device));
    constexpr int CAP = 512;
    constexpr size_t STAGE_MAX = 48u << 10;  // LDS read stage per block
    const uint32_t rw = in.empty() ? 0u : psx_read_words(in[0].p);
    for (size_t k = 0; k < in.size(); ++k) {
        const psx_mutations* m = in[k].m;
        const psx_forest* f = in[k].f;
        if (m->ploidy > 16)
            psx_throw(PSX_EINVAL, "read records hold copy indices below 16 (job %zu: ploidy %u)", k, m->ploidy);
        if (f->n_leaves >= (1u << 26))
            psx_throw(PSX_EINVAL, "read records hold leaf ranks below 2^26 (job %zu: the forest has %u leaves)", k,
                      f->n_leaves);
        if (psx_read_words(in[k].p) != rw)
            psx_throw(PSX_EINVAL, "jobs of one read buffer need one read size (job %zu: %u words per read, job 0: %u)",
                      k, psx_read_words(in[k].p), rw);
    }
    std::unique_ptr<seq_batch, void (*)(seq_batch*)> B(batch_prepare(ctx, in.data(), in.size(), 1), batch_free);
    seq_batch& P = *B;
    cn_rows cn;
    const bool with_cn = cn.prepare(P);
    // presence flags per (output sample, mutation): the read kernel stages
    // only the SNVs a read of the sample can carry. Kernels index a job's
    // flags from its own base (so * n_mut + m), so each job gets the slice
    // that starts at its first (job, output sample) entry.
    dbuf d_present;
    d_present.alloc(P.pres_total);
    std::vector<uint64_t> pres_base(P.jobs.size(), 0);
    for (const pres_dev& q : P.pres)
        if (q.so == 0) pres_base[q.job] = q.begin;
    job_dev* hj = (job_dev*)(P.sb.host.data() + P.o_jobs);  // the uploaded descriptors
    for (size_t k = 0; k < P.jobs.size(); ++k) hj[k].present = d_present.as<uint8_t>() + pres_base[k];
    const seq_lane lane = ctx_lane(ctx, batch_desc_bytes(&P));
    batch_upload(&P, lane);
    if (with_cn) cn.finish(ctx, P, lane);
    if (P.n_reads > capacity_reads)
        psx_throw(PSX_EINVAL, "the read buffer holds %llu reads, the run simulates %llu",
                  (unsigned long long)capacity_reads, (unsigned long long)P.n_reads);
    if (job_first) {
        job_first[0] = 0;
        for (size_t k = 0; k < P.job_reads.size(); ++k) job_first[k + 1] = job_first[k] + P.job_reads[k];
    }
    seq_launch L = launch_desc(P, lane.desc, lane.queue);
    PSX_HIP(hipMemsetAsync(lane.queue, 0, 8, ctx->stream));
    if (P.pres_total) {
        const pres_dev* ps = (const pres_dev*)(lane.desc + P.o_pres);
        const uint32_t pgrid = (uint32_t)((P.pres_total + 255) / 256);
        hipLaunchKernelGGL(k_reads_present, dim3(pgrid), dim3(256), 0, ctx->stream, L.jobs, ps,
                           (uint32_t)P.pres.size(), P.pres_total);
    }
    L.reads_out = d_reads;
    L.read_words = rw;
    L.out_vec4 = (L.read_words % 4 == 0 && ((uintptr_t)d_reads & 15) == 0) ? 1u : 0u;
    // 64 rows of read_words words per wave staged in LDS; the block is the
    // largest of 512 and 256 threads whose stage fits STAGE_MAX (more waves
    // per CU share one copy of the block's window, reference and tables:
    // 512 threads run 6 waves per SIMD at 150 bp, 256 threads 4 or 5)
    const size_t wave_stage = (size_t)64 * L.read_words * 4u;
    if (P.n_items == 0) PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (P.n_items) {
        const size_t items = P.n_items;
        if (with_cn) {  // one instantiation with the copy-number path (any read size)
            if (8 * wave_stage <= STAGE_MAX)
                launch_reads<512, CAP, true, 0, true>(ctx, L, items, 8 * wave_stage);
            else if (4 * wave_stage <= STAGE_MAX)
                launch_reads<256, CAP, true, 0, true>(ctx, L, items, 4 * wave_stage);
            else
                launch_reads<256, CAP, false, 0, true>(ctx, L, items, 0);
        } else if (L.out_vec4 && L.read_words == 12) {  // 145-160 bp reads
            launch_reads<512, CAP, true, 10>(ctx, L, items, 8 * wave_stage);
        } else if (8 * wave_stage <= STAGE_MAX) {
            launch_reads<512, CAP, true, 0>(ctx, L, items, 8 * wave_stage);
        } else if (4 * wave_stage <= STAGE_MAX) {
            launch_reads<256, CAP, true, 0>(ctx, L, items, 4 * wave_stage);
        } else {
            launch_reads<256, CAP, false, 0>(ctx, L, items, 0);
        }
    }
    PSX_HIP(hipGetLastError());
    PSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    PSX_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_kernel_ms = ms;
    if (n_reads) *n_reads = P.n_reads;
}

}  // namespace

// Reads of each (output sample, chromosome) segment of one job, [n_out][n_chr]
// (0 for chromosomes left out or shorter than a read): the closed form, or
// with copy-number events the device count of kept templates.
std::vector<uint64_t> psxh::segment_reads(psx_ctx* ctx, const job_in& in) {
    std::unique_ptr<seq_batch, void (*)(seq_batch*)> B(batch_prepare(ctx, &in, 1, 1), batch_free);
    seq_batch& P = *B;
    cn_rows cn;
    if (cn.prepare(P)) {
        const seq_lane lane = ctx_lane(ctx, batch_desc_bytes(&P));
        batch_upload(&P, lane);
        cn.finish(ctx, P, lane);
    }
    const seg_dev* hs = (const seg_dev*)(P.sb.host.data() + P.o_segs);
    const job_dev* hj = (const job_dev*)(P.sb.host.data() + P.o_jobs);
    const uint32_t n_chr = in.g->n_chr;
    std::vector<uint64_t> out((size_t)hj[0].n_out * n_chr, 0);
    for (size_t k = 0; k < P.segs.size(); ++k) {
        const seg_dev& S = hs[k];
        out[(size_t)S.so * n_chr + S.chr] = cn.seg_reads.empty() ? S.T * (hj[0].pe ? 2u : 1u) : cn.seg_reads[k];
    }
    return out;
}

extern "C" psx_status psx_seq_num_reads_table(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                              const psx_mutations* m, const psx_seq_params* p, uint64_t seed,
                                              uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p || !n_reads) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        const std::vector<uint64_t> r = psxh::segment_reads(ctx, job_in{g, f, m, p, seed, nullptr, nullptr, nullptr});
        uint64_t n = 0;
        for (uint64_t x : r) n += x;
        *n_reads = n;
    });
}

extern "C" psx_status psx_simulate_seq_reads(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                             const psx_mutations* m, const psx_seq_params* p, uint64_t seed,
                                             uint32_t* d_reads, uint64_t capacity_reads, uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p || !d_reads) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        run_reads(ctx, {job_in{g, f, m, p, seed, nullptr, nullptr, nullptr}}, d_reads, capacity_reads, nullptr,
                  n_reads);
    });
}

extern "C" psx_status psx_simulate_seq_reads_batch(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                                   uint32_t* d_reads, uint64_t capacity_reads, uint64_t* job_first,
                                                   uint64_t* n_reads) {
    if (!ctx || (!jobs && n_jobs) || (!d_reads && n_jobs)) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        std::vector<job_in> in;
        for (uint32_t k = 0; k < n_jobs; ++k) {
            const psx_seq_job& J = jobs[k];
            if (!J.genome || !J.forest || !J.mutations) psx_throw(PSX_EINVAL, "job %u is incomplete", k);
            in.push_back(job_in{J.genome, J.forest, J.mutations, &J.params, J.seed, nullptr, nullptr, nullptr});
        }
        if (in.empty()) {
            if (job_first) job_first[0] = 0;
            if (n_reads) *n_reads = 0;
            return;
        }
        run_reads(ctx, in, d_reads, capacity_reads, job_first, n_reads);
    });
}

extern "C" psx_status psx_simulate_seq_batch(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                             const psx_launch_config* cfg, uint64_t* n_reads) {
    if (!ctx || (!jobs && n_jobs)) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        const std::vector<job_in> in = jobs_of(jobs, n_jobs);
        psx_launch_config c = cfg ? *cfg : (ctx->tuned ? ctx->tuned_cfg : default_config());
        if (!cfg) c.use_graph = 1;
        run_result r = in.empty() ? run_result{0, 0.0} : run_count(ctx, in, c);
        if (n_reads) *n_reads = r.n_reads;
    });
}

// A batched launch prepared once: descriptors in its own device buffer and,
// with use_graph, the init and count kernels captured in a hipGraph that
// every run replays.
struct psx_seq_plan {
    psx_ctx* ctx = nullptr;
    std::unique_ptr<seq_batch, void (*)(seq_batch*)> batch{nullptr, batch_free};
    psx_launch_config cfg{};
    dbuf desc, queue;
    seq_lane lane;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" psx_status psx_seq_plan_create(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                          const psx_launch_config* cfg, psx_seq_plan** out) {
    if (!ctx || !jobs || !n_jobs || !out) return PSX_EINVAL;
    *out = nullptr;
    auto P = std::make_unique<psx_seq_plan>();
    psx_status st = psx_guard(ctx, [&] {
        const std::vector<job_in> in = jobs_of(jobs, n_jobs);
        psx_launch_config c = cfg ? *cfg : (ctx->tuned ? ctx->tuned_cfg : default_config());
        if (!cfg) c.use_graph = 1;
        check_config(c);
        PSX_HIP(hipSetDevice(ctx->device));
        P->ctx = ctx;
        P->cfg = c;
        P->batch.reset(batch_prepare(ctx, in.data(), in.size(), c.strata_per_item));
        P->desc.alloc(batch_desc_bytes(P->batch.get()));
        P->queue.alloc(256);
        P->lane.n_cu = ctx->n_cu;
        P->lane.stream = ctx->stream;
        P->lane.ev0 = ctx->ev0;
        P->lane.ev1 = ctx->ev1;
        P->lane.desc = (uint8_t*)P->desc.p;
        P->lane.desc_bytes = P->desc.bytes;
        P->lane.queue = P->queue.as<unsigned long long>();
        batch_upload(P->batch.get(), P->lane);
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        if (c.use_graph) {
            PSX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            batch_launch(P->batch.get(), P->lane, c, false);
            PSX_HIP(hipStreamEndCapture(ctx->stream, &P->graph));
            PSX_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
        }
    });
    if (st != PSX_OK) {
        psx_seq_plan_free(P.release());
        return st;
    }
    *out = P.release();
    return PSX_OK;
}

extern "C" psx_status psx_seq_plan_run(psx_seq_plan* P, uint64_t* n_reads) {
    if (!P || !P->ctx) return PSX_EINVAL;
    psx_ctx* ctx = P->ctx;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        if (P->exec) {
            PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
            PSX_HIP(hipGraphLaunch(P->exec, ctx->stream));
            PSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        } else {
            batch_launch(P->batch.get(), P->lane, P->cfg, true);
        }
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        float ms = 0.f;
        PSX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_kernel_ms = ms;
        if (n_reads) {
            uint64_t n = 0;
            for (uint64_t r : batch_job_reads(P->batch.get(), P->lane)) n += r;
            *n_reads = n;
        }
    });
}

extern "C" void psx_seq_plan_free(psx_seq_plan* P) {
    if (!P) return;
    if (P->ctx) (void)hipSetDevice(P->ctx->device);
    if (P->exec) (void)hipGraphExecDestroy(P->exec);
    if (P->graph) (void)hipGraphDestroy(P->graph);
    delete P;
}

extern "C" psx_status psx_autotune_seq(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                       psx_launch_config* best, double* best_ms) {
    if (!ctx || !jobs || !n_jobs || !best) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        const std::vector<job_in> in = jobs_of(jobs, n_jobs);
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
                        const psx_launch_config c{b, cap, spi, bpc, 0};
                        psx_check_cancel(ctx);
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
