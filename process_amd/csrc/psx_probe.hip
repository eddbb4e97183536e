// Measurement kernels (no counterpart in the reference): the vector-ALU issue
// ceiling against which the count-only sequencing kernel is reported. That
// kernel moves well under a byte of HBM per read, so its bound is the rate at
// which the compute units issue vector instructions, and the rate reachable
// depends on the instruction mix; these kernels measure it on the running GPU
// with no memory traffic in the loop:
//
//   PSX_PROBE_PHILOX   every lane draws Philox4x32-10 blocks (the same
//                      psx_draw_u as the sequencing kernels) and folds them
//                      with three-input xors: the template and error draws
//                      of the count kernel, alone;
//   PSX_PROBE_INT      every lane runs eight independent chains of 32-bit
//                      integer adds and three-input xors: single-cycle
//                      vector instructions with no dependency stalls.
//
// Each lane writes one word at the end so nothing is optimised away. The
// launch fills every compute unit with 256-thread blocks (8 per CU, the
// occupancy of the count kernel's launch), and is timed with HIP events.
#include "psx_internal.h"

namespace {

template <int MODE>
__global__ void __launch_bounds__(256) k_probe_valu(uint64_t seed, uint32_t iters, uint32_t* __restrict__ out) {
    const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
    uint32_t acc = 0;
    if constexpr (MODE == PSX_PROBE_PHILOX) {
        for (uint32_t i = 0; i < iters; ++i) {
            const psx_u4 w = psx_draw_u(seed, tid, i, 0u, PSX_P_TEMPLATE << 24);
            acc = psx_xor3(acc, psx_xor3(w.x, w.y, w.z), w.w);
        }
    } else {
        uint32_t a[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = tid * (2u * k + 1u);
        for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
            for (int k = 0; k < 8; ++k) a[k] = psx_xor3(a[k] + i, a[k] >> 3, 0x9E3779B9u * (k + 1));
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc ^= a[k];
    }
    out[tid] = acc;
}

}  // namespace

extern "C" psx_status psx_probe_valu(psx_ctx* ctx, int mode, uint32_t iters, double* ms, uint64_t* lane_iters) {
    if (!ctx || !ms || (mode != PSX_PROBE_PHILOX && mode != PSX_PROBE_INT)) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        const uint32_t grid = (uint32_t)ctx->n_cu * 8u;
        const size_t bytes = (size_t)grid * 256u * 4u;
        ctx->scratch.reserve(bytes);
        uint32_t* out = ctx->scratch.as<uint32_t>();
        PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        if (mode == PSX_PROBE_PHILOX)
            hipLaunchKernelGGL(k_probe_valu<PSX_PROBE_PHILOX>, dim3(grid), dim3(256), 0, ctx->stream, 0x5EEDull,
                               iters, out);
        else
            hipLaunchKernelGGL(k_probe_valu<PSX_PROBE_INT>, dim3(grid), dim3(256), 0, ctx->stream, 0x5EEDull, iters,
                               out);
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        float t = 0.f;
        PSX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        *ms = t;
        if (lane_iters) *lane_iters = (uint64_t)grid * 256u * iters;
    });
}
