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
//
// And the HBM write ceiling against which the read-materialising kernel is
// reported: stores only, into a caller's buffer,
//
//   PSX_PROBE_STORE_FILL     a grid-stride fill, consecutive lanes writing
//                            consecutive 16-byte words;
//   PSX_PROBE_STORE_CHUNKS   the read kernel's shape: items of 170 chunks of
//                            3 KiB (a stratum's rows) taken from a counter by
//                            512-thread blocks, each wave writing 2 chunks per
//                            round of 16 (3 stores of 1 KiB per chunk), the
//                            chunks on 1024-byte boundaries;
//   PSX_PROBE_STORE_CHUNKS_SHIFTED  the same 16 bytes off those boundaries
//                            (every chunk boundary inside a cache line).
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

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_probe_fill(v4u* __restrict__ out, uint64_t n_u4) {
    const uint64_t i0 = (uint64_t)blockIdx.x * 256u + threadIdx.x, st = (uint64_t)gridDim.x * 256u;
    for (uint64_t i = i0; i < n_u4; i += st) out[i] = v4u{(uint32_t)i, 1u, 2u, 3u};
}

constexpr uint32_t PROBE_IC = 170;     // chunks per item
constexpr uint32_t PROBE_CHUNK = 3072;  // bytes per chunk

__global__ void __launch_bounds__(512) k_probe_chunks(char* __restrict__ base, uint32_t n_items,
                                                      unsigned long long* __restrict__ counter) {
    __shared__ uint32_t s_item;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (;;) {
        if (tid == 0) s_item = (uint32_t)atomicAdd(counter, 1ull);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= n_items) break;
        for (uint32_t r = 0; r * 16u < PROBE_IC; ++r)
            for (uint32_t h = 0; h < 2u; ++h) {
                const uint32_t k = r * 16u + wave * 2u + h;
                if (k >= PROBE_IC) continue;
                v4u* d = reinterpret_cast<v4u*>(base + ((uint64_t)item * PROBE_IC + k) * PROBE_CHUNK);
#pragma unroll
                for (uint32_t q = 0; q < 3u; ++q) d[q * 64u + lane] = v4u{item, k, q, 3u};
            }
    }
}

}  // namespace

extern "C" psx_status psx_probe_store(psx_ctx* ctx, void* out, uint64_t bytes, int mode, double* ms,
                                      uint64_t* written) {
    if (!ctx || !out || !ms || !written || ((uintptr_t)out & 15u) ||
        (mode != PSX_PROBE_STORE_FILL && mode != PSX_PROBE_STORE_CHUNKS && mode != PSX_PROBE_STORE_CHUNKS_SHIFTED))
        return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        uint64_t n = 0;
        if (mode == PSX_PROBE_STORE_FILL) {
            n = bytes / 16u * 16u;
            PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
            hipLaunchKernelGGL(k_probe_fill, dim3((uint32_t)ctx->n_cu), dim3(256), 0, ctx->stream, (v4u*)out, n / 16u);
        } else {
            // chunks from the first 1024-byte boundary of the buffer (+ 16 bytes when shifted)
            const uintptr_t p0 = (uintptr_t)out, b = ((p0 + 1023u) & ~(uintptr_t)1023u) +
                                                       (mode == PSX_PROBE_STORE_CHUNKS_SHIFTED ? 16u : 0u);
            const uint64_t skip = b - p0;
            const uint64_t per_item = (uint64_t)PROBE_IC * PROBE_CHUNK;
            const uint64_t n_items = bytes > skip ? (bytes - skip) / per_item : 0u;
            if (n_items == 0 || n_items > 0xFFFFFFFFull) psx_throw(PSX_EINVAL, "buffer size out of range for the probe");
            n = n_items * per_item;  // skip + n <= bytes
            ctx->scratch.reserve(8);
            unsigned long long* counter = ctx->scratch.as<unsigned long long>();
            PSX_HIP(hipMemsetAsync(counter, 0, 8, ctx->stream));
            PSX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
            hipLaunchKernelGGL(k_probe_chunks, dim3((uint32_t)ctx->n_cu * 3u), dim3(512), 0, ctx->stream, (char*)b,
                               (uint32_t)n_items, counter);
        }
        PSX_HIP(hipGetLastError());
        PSX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        PSX_HIP(hipStreamSynchronize(ctx->stream));
        float t = 0.f;
        PSX_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        *ms = t;
        *written = n;
    });
}

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
