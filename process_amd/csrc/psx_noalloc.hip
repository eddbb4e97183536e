// Count-only sequencing into caller-owned device buffers with no device
// allocation: the launch descriptors and work counters live in a fixed
// file-scope device array instead of buffers the library allocates, for
// callers that cannot allocate (code under HIP graph capture, serving loops
// and harnesses that forbid hipMalloc). No counterpart in the reference.
#include <mutex>

#include "psx_seq_host.h"

using namespace psxh;

// One copy per GPU, held until the process ends. Host code never passes this
// name to a kernel or a copy: it uses the device address looked up for the
// device (g_scratch_ptr).
static __device__ __align__(256) uint8_t scratch[PSX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_DEVICES = 64;
constexpr size_t SCRATCH_QUEUE = 256;  // work counters first, then the descriptors

std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[SCRATCH_DEVICES] = {};

}  // namespace

uint8_t* psx_noalloc_scratch_init(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES)
        psx_throw(PSX_EINVAL, "device %d is beyond the %d devices of the no-allocation scratch", device,
                  SCRATCH_DEVICES);
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        PSX_HIP(hipGetDevice(&cur));
        if (cur != device)
            psx_throw(PSX_EHIP, "the no-allocation scratch of device %d is looked up with device %d current", device,
                      cur);
        void* ptr = nullptr;
        PSX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) psx_throw(PSX_EHIP, "no device address for the no-allocation scratch of device %d", device);
        g_scratch_ptr[device] = (uint8_t*)ptr;
    }
    return g_scratch_ptr[device];
}

extern "C" psx_status psx_simulate_seq_device_noalloc(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                                      const psx_mutations* m, const psx_seq_params* p, uint64_t seed,
                                                      int32_t* d_alt, int32_t* d_depth, uint8_t* d_present,
                                                      uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p || !d_alt || !d_depth || !d_present) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        if (!ctx->noalloc_scratch) psx_throw(PSX_EHIP, "the context has no no-allocation scratch");
        psx_launch_config c = ctx->tuned ? ctx->tuned_cfg : default_config();
        c.use_graph = 0;
        check_config(c);
        PSX_HIP(hipSetDevice(ctx->device));
        const job_in in{g, f, m, p, seed, d_alt, d_depth, d_present};
        seq_batch* P = batch_prepare(ctx, &in, 1, c.strata_per_item, true);
        const size_t need = batch_desc_bytes(P);
        if (need > PSX_NOALLOC_SCRATCH_BYTES - SCRATCH_QUEUE) {
            batch_free(P);
            psx_throw(PSX_EINVAL, "the call needs %zu bytes of launch descriptors; the no-allocation scratch holds %zu",
                      need, (size_t)PSX_NOALLOC_SCRATCH_BYTES - SCRATCH_QUEUE);
        }
        seq_lane lane;
        lane.n_cu = ctx->n_cu;
        lane.stream = ctx->stream;
        lane.ev0 = ctx->ev0;
        lane.ev1 = ctx->ev1;
        lane.queue = (unsigned long long*)ctx->noalloc_scratch;  // zeroed by the init kernel of every launch
        lane.desc = ctx->noalloc_scratch + SCRATCH_QUEUE;
        lane.desc_bytes = PSX_NOALLOC_SCRATCH_BYTES - SCRATCH_QUEUE;
        const run_result r = run_prepared(ctx, P, lane, c);
        if (n_reads) *n_reads = r.n_reads;
    });
}
