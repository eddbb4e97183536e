// Counter-based random draws shared by the gfx950 kernels and the host code.
//
// Philox4x32-10 follows Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
// as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64 multiplies with the
// published multipliers and Weyl key increments. Every random decision of the
// library is a pure function of (key, 128-bit counter), so a result does not
// depend on thread mapping, launch shape, batching or co-scheduled work.
//
// Samplers are integer-only (no libm on the device):
//   * uniform in [0, n):   Lemire multiply-shift, floor(u * n / 2^32) (or the
//                          64-bit variant); no rejection, so the relative bias
//                          is at most n / 2^32 (n / 2^64 for the wide form);
//   * categorical / Poisson / binomial: inversion of a cumulative table
//                          frozen on the host to integers T[k] = floor(cdf_k *
//                          2^32), T[last] = 2^32; the draw is the first k with
//                          u < T[k].
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PSX_HD __host__ __device__ __forceinline__
#else
#define PSX_HD static inline
#endif

// Philox counter purposes: the high byte of counter word 3.
enum psx_purpose : uint32_t {
    PSX_P_REF = 0x01u,       // synthetic reference bases
    PSX_P_TEMPLATE = 0x10u,  // read template: start, tumour/normal, source copy, insert
    PSX_P_ERROR = 0x11u,     // sequencing errors of a template (both mates)
    PSX_P_QUALITY = 0x12u,   // quality scores of a written read (random quality model)
    PSX_P_EDGE = 0x20u,      // per-edge passenger counts
    PSX_P_SNV = 0x21u,       // SNV class and allele
    PSX_P_SITE = 0x22u,      // SNV site: index into the context index list of the class
    PSX_P_INDEL = 0x23u,     // indel type, length, allele, position
    PSX_P_IDCLS = 0x24u,     // indel signature, ID class and allele (ID signatures)
    PSX_P_DRIVER = 0x25u,    // driver allele
    PSX_P_IDSITE = 0x26u,    // indel site search (ID signatures)
    PSX_P_PCNA = 0x27u,      // passenger CNA: catalogue region and allele
    PSX_P_FOREST = 0x30u,    // synthetic birth-death forest
    PSX_P_PICK = 0x31u,      // synthetic sampling of cells
};

struct psx_u4 {
    uint32_t x, y, z, w;
};

PSX_HD uint32_t psx_mulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

PSX_HD uint64_t psx_mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * (unsigned __int128)b) >> 64);
#endif
}

// a ^ b ^ c (one v_bitop3_b32 on gfx950; truth table 0x96 is the three-input xor)
PSX_HD uint32_t psx_xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

PSX_HD psx_u4 psx_philox(psx_u4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 multiply each (a single v_mad_u64_u32 on gfx950)
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        c = psx_u4{psx_xor3((uint32_t)(p1 >> 32), c.y, k0), (uint32_t)p1, psx_xor3((uint32_t)(p0 >> 32), c.w, k1),
                   (uint32_t)p0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

PSX_HD psx_u4 psx_draw(uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    return psx_philox(psx_u4{c0, c1, c2, c3}, (uint32_t)key, (uint32_t)(key >> 32));
}

#if defined(__HIPCC__)
// psx_draw for a key that is uniform across the wavefront (a launch seed).
// The empty asm keeps the key in scalar registers and stops the compiler from
// hoisting the ten round keys out of the caller's loops: kept live, they
// exhaust the scalar registers and are spilled to vector lanes, and every
// round then pays vector read-lane instructions; recomputed, each round key
// costs one scalar add.
__device__ __forceinline__ psx_u4 psx_draw_u(uint64_t key, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3) {
    uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+s"(k0), "+s"(k1));
#endif
    return psx_philox(psx_u4{c0, c1, c2, c3}, k0, k1);
}
#endif

PSX_HD uint32_t psx_word(const psx_u4& v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// floor(u * n / 2^32): uniform in [0, n) for n <= 2^32.
PSX_HD uint32_t psx_uniform32(uint32_t u, uint32_t n) { return psx_mulhi32(u, n); }

// floor(((hi << 32) | lo) * n / 2^64): uniform in [0, n).
PSX_HD uint64_t psx_uniform64(uint32_t hi, uint32_t lo, uint64_t n) {
    return psx_mulhi64(((uint64_t)hi << 32) | lo, n);
}

// Compact form of an inversion table for the kernels: entries that can never
// be drawn are dropped (leading T[k] == 0; everything after the first
// T[k] == 2^32), so every kept threshold but the last fits in 32 bits, and a
// guide table of PSX_GUIDE entries, g[i] = first k with T[k] > i * 2^(32 -
// PSX_GUIDE_BITS), starts the scan. The draw equals psx_cdf_pick on the full
// table whatever the guide's size.
#define PSX_GUIDE_BITS 11u
#define PSX_GUIDE (1u << PSX_GUIDE_BITS)
PSX_HD uint32_t psx_guided_pick(const uint32_t* T32, const uint16_t* guide, uint32_t n, uint32_t u) {
    uint32_t k = guide[u >> (32u - PSX_GUIDE_BITS)];
    while (k + 1 < n && u >= T32[k]) ++k;
    return k;
}

// First k in [0, n) with u < T[k]; T is non-decreasing with T[n-1] = 2^32.
PSX_HD uint32_t psx_cdf_pick(const uint64_t* T, uint32_t n, uint32_t u) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)u < T[mid]) hi = mid; else lo = mid + 1;
    }
    return lo;
}
