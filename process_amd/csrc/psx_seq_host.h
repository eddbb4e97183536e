// Host-side interface of the sequencing launches, shared by the one-call
// entry points (psx_seq.hip) and the serving runtime (psx_serve.cpp).
// Not part of the C ABI.
#pragma once
#include <functional>

#include "psx_internal.h"

namespace psxh {

// One sequencing job: inputs, parameters and device output arrays
// (n_out_samples x n_mutations each).
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

// Where a launch runs: its stream, the events around the count kernel, a
// device descriptor buffer of desc_bytes and an 8-byte work counter.
struct seq_lane {
    int n_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint8_t* desc = nullptr;
    size_t desc_bytes = 0;
    unsigned long long* queue = nullptr;
};

struct seq_batch;  // descriptors of one launch over one or more jobs

// Validates the jobs (psx_error on the first invalid one) and builds their
// descriptors on the host. Builds a mutation table's stratum index for a
// genome it was not indexed for yet (allocates and synchronises ctx's stream)
// unless no_alloc, which refuses such a job instead.
seq_batch* batch_prepare(psx_ctx* ctx, const job_in* jobs, size_t n_jobs, uint32_t strata_per_item,
                         bool no_alloc = false);
void batch_free(seq_batch* b);
size_t batch_desc_bytes(const seq_batch* b);
uint64_t batch_reads(const seq_batch* b);
size_t batch_items(const seq_batch* b);
// Reads simulated per job once the launch has finished (jobs with
// copy-number events: the device count of accepted templates).
std::vector<uint64_t> batch_job_reads(const seq_batch* b, const seq_lane& lane);
// Copies the descriptors to lane.desc (which must hold batch_desc_bytes).
void batch_upload(seq_batch* b, const seq_lane& lane);
// Enqueues the init kernel and the count kernel on lane.stream; with
// timed != 0 the count kernel is bracketed by lane.ev0 / lane.ev1.
void batch_launch(const seq_batch* b, const seq_lane& lane, const psx_launch_config& cfg, bool timed);

// Reads of each (output sample, chromosome) segment of one job, [n_out][n_chr]
// in output order (0 where a chromosome is left out or shorter than a read):
// the closed form, or with copy-number events the kept templates counted on
// ctx's stream (synchronises it).
std::vector<uint64_t> segment_reads(psx_ctx* ctx, const job_in& in);

struct run_result {
    uint64_t n_reads;
    double kernel_ms;
};
// Takes ownership of a prepared batch: uploads its descriptors to the lane,
// runs the init and count kernels (captured in a hipGraph and replayed once
// with cfg.use_graph), waits, and returns the reads simulated and the
// HIP-event time of the count kernel (of the graph launch when captured),
// also kept in ctx->last_kernel_ms. `during`, when given, runs on the host
// while the kernels run (after the launch, before the wait).
run_result run_prepared(psx_ctx* ctx, seq_batch* batch, const seq_lane& lane, const psx_launch_config& cfg,
                        const std::function<void()>& during = {});

uint32_t out_samples(const psx_forest* f, const psx_seq_params* p);
psx_launch_config default_config();
void check_config(const psx_launch_config& c);

}  // namespace psxh
