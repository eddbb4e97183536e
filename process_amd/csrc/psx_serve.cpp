// Serving runtime: continuous batching of count-only sequencing requests on
// one GPU. No counterpart in the reference (simulate_seq runs one call at a
// time, src/seq_simulation.cpp:517-601); the launches it makes are the
// batched launches of psx_seq.hip, so every result equals psx_simulate_seq
// on the same inputs.
//
// Design. One device allocation (the HBM budget) holds the outputs of the
// admitted requests; a host-side first-fit arena hands out ranges of it, so
// admission never calls hipMalloc or hipFree while kernels run. Each launch
// slot ("lane") is a worker thread with its own stream, events, work counter
// and descriptor buffer. A worker that is free takes every queued request
// that fits in the arena (arrival order, up to max_batch) into one
// persistent launch, copies the results to the callers' host arrays, returns
// the memory and wakes the waiters; lanes run concurrently, so new requests
// start while earlier launches are still on the GPU.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "psx_arena.h"
#include "psx_seq_host.h"

using namespace psxh;

namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct srv_req {
    uint64_t ticket = 0;
    psx_request r{};
    std::vector<uint32_t> chroms, samples;
    uint64_t n_elems = 0;      // n_out_samples x n_mutations
    size_t out_bytes = 0;
    size_t off = arena::NONE;  // arena block of the outputs
    clk::time_point t_submit, t_launch, t_done;
    psx_request_stats st{};
    bool done = false;
    std::string err;
};

// alt | depth | present, each 256-byte aligned
size_t out_layout(uint64_t n, size_t* o_depth, size_t* o_present) {
    const size_t a = arena::round(n * 4);
    *o_depth = a;
    *o_present = 2 * a;
    return 2 * a + arena::round(n);
}

struct lane_state {
    seq_lane lane;
    dbuf desc, queue;
    std::vector<dbuf> retired;  // outgrown descriptor buffers, freed at destroy
    std::thread th;
};

}  // namespace

struct psx_server {
    psx_ctx* ctx = nullptr;
    psx_server_config cfg{};
    dbuf pool;
    arena ar;
    std::mutex mu;
    std::condition_variable cv;  // work queued, memory returned, request finished, stop
    std::deque<srv_req*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<srv_req>> reqs;
    std::vector<std::unique_ptr<lane_state>> lanes;
    std::mutex prep_mu;          // descriptor preparation may re-index a shared mutation table
    std::vector<clk::time_point> lane_start;  // start of each lane's running launch
    std::vector<char> lane_busy;
    bool stop = false;
    uint64_t next_ticket = 1;
    psx_server_totals tot{};
    std::string err;
};

namespace {

void run_batch(psx_server* s, lane_state& L, std::vector<srv_req*>& batch) {
    uint8_t* base = (uint8_t*)s->pool.p;
    std::vector<job_in> jobs;
    for (srv_req* q : batch) {
        size_t od, op;
        out_layout(q->n_elems, &od, &op);
        uint8_t* o = base + q->off;
        jobs.push_back(job_in{q->r.genome, q->r.forest, q->r.mutations, &q->r.params, q->r.seed, (int32_t*)o,
                              (int32_t*)(o + od), o + op});
    }
    std::unique_ptr<seq_batch, void (*)(seq_batch*)> B(nullptr, batch_free);
    {
        std::lock_guard<std::mutex> g(s->prep_mu);
        B.reset(batch_prepare(s->ctx, jobs.data(), jobs.size(), s->cfg.launch.strata_per_item));
    }
    const size_t need = batch_desc_bytes(B.get());
    if (L.desc.bytes < need) {
        // a buffer still read by an earlier launch of this lane is kept until destroy
        L.retired.push_back(std::move(L.desc));
        L.desc.alloc(need * 2);
        L.lane.desc = (uint8_t*)L.desc.p;
        L.lane.desc_bytes = L.desc.bytes;
    }
    batch_upload(B.get(), L.lane);
    const clk::time_point t_launch = clk::now();
    for (srv_req* q : batch) q->t_launch = t_launch;
    batch_launch(B.get(), L.lane, s->cfg.launch, true);
    for (srv_req* q : batch) {
        if (!q->n_elems) continue;
        size_t od, op;
        out_layout(q->n_elems, &od, &op);
        uint8_t* o = base + q->off;
        if (q->r.alt) PSX_HIP(hipMemcpyAsync(q->r.alt, o, q->n_elems * 4, hipMemcpyDeviceToHost, L.lane.stream));
        if (q->r.depth)
            PSX_HIP(hipMemcpyAsync(q->r.depth, o + od, q->n_elems * 4, hipMemcpyDeviceToHost, L.lane.stream));
        if (q->r.present)
            PSX_HIP(hipMemcpyAsync(q->r.present, o + op, q->n_elems, hipMemcpyDeviceToHost, L.lane.stream));
    }
    PSX_HIP(hipStreamSynchronize(L.lane.stream));
    const std::vector<uint64_t> reads = batch_job_reads(B.get(), L.lane);
    for (size_t k = 0; k < batch.size(); ++k) batch[k]->st.n_reads = reads[k];
}

void worker(psx_server* s, uint32_t k) {
    lane_state& L = *s->lanes[k];
    const bool dev_ok = hipSetDevice(s->ctx->device) == hipSuccess;
    std::unique_lock<std::mutex> lk(s->mu);
    const auto overlap = std::chrono::microseconds(s->cfg.overlap_after_us);
    for (;;) {
        s->cv.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        if (s->stop) break;
        if (s->cfg.overlap_after_us && s->pending.size() < s->cfg.max_batch) {
            // another lane's young launch: wait until it has run overlap_after_us (or ends)
            bool young = false;
            clk::time_point until{};
            for (size_t j = 0; j < s->lanes.size(); ++j)
                if (s->lane_busy[j] && (!young || s->lane_start[j] + overlap < until)) {
                    young = true;
                    until = s->lane_start[j] + overlap;
                }
            if (young && clk::now() < until) {
                s->cv.wait_until(lk, until);
                continue;
            }
        }
        std::vector<srv_req*> batch;
        admit_fifo(s->pending, s->ar, s->cfg.max_batch, batch, [](srv_req* q) { return q->out_bytes; },
                   [](srv_req* q, size_t off) { q->off = off; });
        if (batch.empty()) {
            // the head request waits for memory held by running launches
            s->cv.wait(lk);
            continue;
        }
        s->tot.launches++;
        s->lane_busy[k] = 1;
        s->lane_start[k] = clk::now();
        lk.unlock();
        psx_status st = PSX_OK;
        std::string msg;
        try {
            if (!dev_ok) psx_throw(PSX_EHIP, "hipSetDevice failed in a serving lane");
            run_batch(s, L, batch);
        } catch (const psx_error& e) {
            st = e.code;
            msg = e.msg;
        } catch (const std::exception& e) {
            st = PSX_EINVAL;
            msg = e.what();
        }
        const clk::time_point t_done = clk::now();
        lk.lock();
        s->lane_busy[k] = 0;
        for (srv_req* q : batch) {
            s->ar.give(q->off);
            q->off = arena::NONE;
            q->t_done = t_done;
            q->st.status = st;
            q->err = msg;
            q->st.batch_size = (uint32_t)batch.size();
            q->st.lane = k;
            q->st.queued_ms = ms_between(q->t_submit, q->t_launch);
            q->st.latency_ms = ms_between(q->t_submit, t_done);
            q->done = true;
            if (st == PSX_OK) {
                s->tot.completed++;
                s->tot.reads += q->st.n_reads;
            } else {
                s->tot.failed++;
            }
        }
        s->cv.notify_all();
    }
}

}  // namespace

extern "C" psx_status psx_server_create(psx_ctx* ctx, const psx_server_config* cfg, psx_server** out) {
    if (!ctx || !cfg || !out) return PSX_EINVAL;
    *out = nullptr;
    auto s = std::make_unique<psx_server>();
    psx_status st = psx_guard(ctx, [&] {
        if (cfg->max_inflight < 1 || cfg->max_inflight > 8)
            psx_throw(PSX_EINVAL, "max_inflight must be in [1, 8] (got %u)", cfg->max_inflight);
        if (cfg->max_batch < 1) psx_throw(PSX_EINVAL, "max_batch must be at least 1");
        if (cfg->hbm_budget < (1u << 20)) psx_throw(PSX_EINVAL, "the HBM budget must be at least 1 MiB");
        check_config(cfg->launch);
        PSX_HIP(hipSetDevice(ctx->device));
        s->ctx = ctx;
        s->cfg = *cfg;
        s->cfg.launch.use_graph = 0;
        s->pool.alloc(cfg->hbm_budget);
        s->ar.reset(cfg->hbm_budget);
        s->tot.budget_bytes = s->ar.total();
        s->lane_start.assign(cfg->max_inflight, clk::time_point{});
        s->lane_busy.assign(cfg->max_inflight, 0);
        for (uint32_t k = 0; k < cfg->max_inflight; ++k) {
            auto L = std::make_unique<lane_state>();
            L->lane.n_cu = ctx->n_cu;
            PSX_HIP(hipStreamCreateWithFlags(&L->lane.stream, hipStreamNonBlocking));
            PSX_HIP(hipEventCreate(&L->lane.ev0));
            PSX_HIP(hipEventCreate(&L->lane.ev1));
            L->queue.alloc(256);
            L->desc.alloc(4u << 20);
            L->lane.queue = L->queue.as<unsigned long long>();
            L->lane.desc = (uint8_t*)L->desc.p;
            L->lane.desc_bytes = L->desc.bytes;
            s->lanes.push_back(std::move(L));
        }
        for (uint32_t k = 0; k < cfg->max_inflight; ++k) s->lanes[k]->th = std::thread(worker, s.get(), k);
    });
    if (st != PSX_OK) {
        psx_server_destroy(s.release());
        return st;
    }
    *out = s.release();
    return PSX_OK;
}

extern "C" psx_status psx_server_submit(psx_server* s, const psx_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return PSX_EINVAL;
    auto q = std::make_unique<srv_req>();
    std::string msg;
    psx_status st = psx_guard(nullptr, [&] {
        if (!r->genome || !r->forest || !r->mutations) psx_throw(PSX_EINVAL, "the request is incomplete");
        q->r = *r;
        psx_seq_params& p = q->r.params;
        if (p.n_chromosomes) {
            if (!p.chromosomes) psx_throw(PSX_EINVAL, "chromosome list missing");
            q->chroms.assign(p.chromosomes, p.chromosomes + p.n_chromosomes);
            p.chromosomes = q->chroms.data();
        }
        if (p.samples) {
            q->samples.assign(p.samples, p.samples + p.n_samples);
            q->samples.push_back(0);  // keeps the pointer non-NULL for an empty list
            p.samples = q->samples.data();
        }
        q->n_elems = (uint64_t)out_samples(r->forest, &p) * r->mutations->n;
        size_t od, op;
        q->out_bytes = out_layout(q->n_elems, &od, &op);
        if (q->out_bytes > s->ar.total())
            psx_throw(PSX_ENOMEM_DEVICE, "the request needs %zu bytes of outputs, the serving budget is %zu",
                      q->out_bytes, s->ar.total());
        const job_in in{r->genome, r->forest, r->mutations, &p, r->seed, nullptr, nullptr, nullptr};
        std::lock_guard<std::mutex> g(s->prep_mu);
        seq_batch* B = batch_prepare(s->ctx, &in, 1, 1);
        q->st.n_reads = batch_reads(B);
        batch_free(B);
    });
    std::lock_guard<std::mutex> lk(s->mu);
    if (st != PSX_OK) {
        s->err = psx_tls_err;
        return st;
    }
    if (s->stop) {
        s->err = "the server is shutting down";
        return PSX_ECANCELLED;
    }
    q->ticket = s->next_ticket++;
    q->t_submit = clk::now();
    *ticket = q->ticket;
    s->pending.push_back(q.get());
    s->reqs[q->ticket] = std::move(q);
    s->tot.submitted++;
    s->cv.notify_all();
    return PSX_OK;
}

extern "C" psx_status psx_server_wait(psx_server* s, uint64_t ticket, double timeout_ms, psx_request_stats* st) {
    if (!s) return PSX_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
        s->err = "unknown ticket " + std::to_string(ticket);
        return PSX_EINVAL;
    }
    srv_req* q = it->second.get();
    auto ready = [&] { return q->done; };
    if (timeout_ms < 0) {
        s->cv.wait(lk, ready);
    } else if (!s->cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
        s->err = "timed out waiting for ticket " + std::to_string(ticket);
        return PSX_ECANCELLED;
    }
    if (st) *st = q->st;
    const psx_status rc = q->st.status;
    if (rc != PSX_OK) s->err = q->err;
    s->reqs.erase(it);
    return rc;
}

extern "C" psx_status psx_server_totals_get(psx_server* s, psx_server_totals* t) {
    if (!s || !t) return PSX_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *t = s->tot;
    t->peak_bytes = s->ar.peak();
    return PSX_OK;
}

extern "C" const char* psx_server_last_error(psx_server* s) { return s ? s->err.c_str() : psx_tls_err.c_str(); }

extern "C" void psx_server_destroy(psx_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (srv_req* q : s->pending) {
            q->st.status = PSX_ECANCELLED;
            q->err = "cancelled: the server was destroyed";
            q->done = true;
        }
        s->pending.clear();
        s->cv.notify_all();
    }
    for (auto& L : s->lanes)
        if (L->th.joinable()) L->th.join();
    if (s->ctx) (void)hipSetDevice(s->ctx->device);
    for (auto& L : s->lanes) {
        if (L->lane.stream) (void)hipStreamSynchronize(L->lane.stream);
        if (L->lane.ev0) (void)hipEventDestroy(L->lane.ev0);
        if (L->lane.ev1) (void)hipEventDestroy(L->lane.ev1);
        if (L->lane.stream) (void)hipStreamDestroy(L->lane.stream);
    }
    delete s;
}
