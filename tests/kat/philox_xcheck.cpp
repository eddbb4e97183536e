// Host-side cross-check of Philox4x32-10: for each input line
// "c0 c1 c2 c3 key" prints the four output words of (a) AMD's open-source
// rocRAND engine seeded so that its counter equals (c0, c1, c2, c3), and
// (b) the psx_rng.h implementation used by the gfx950 kernels.
#include <cstdio>
#include <cstdint>

#include <rocrand/rocrand_philox4x32_10.h>

#include "psx_rng.h"

int main() {
    unsigned long long c0, c1, c2, c3, key;
    while (std::scanf("%llu %llu %llu %llu %llu", &c0, &c1, &c2, &c3, &key) == 5) {
        const unsigned long long subsequence = c2 | (c3 << 32);
        const unsigned long long offset = 4ull * (c0 | (c1 << 32));
        rocrand_device::philox4x32_10_engine eng(key, subsequence, offset);
        unsigned int r[4];
        for (int i = 0; i < 4; ++i) r[i] = eng();
        const psx_u4 p = psx_draw(key, (uint32_t)c0, (uint32_t)c1, (uint32_t)c2, (uint32_t)c3);
        std::printf("%u %u %u %u %u %u %u %u\n", r[0], r[1], r[2], r[3], p.x, p.y, p.z, p.w);
    }
    return 0;
}
