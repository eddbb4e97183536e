// COSMIC indel (ID83) classes of reference sites, shared by the placement
// kernels and the kernels that build a genome's indel index.
//
// For a position q of a chromosome b[0..len) (the indel's anchor base is
// b[q - 1]) and a unit length u in 1..5, U = b[q..q+u) and q is a run start
// when b[q-u..q) != U (q >= u, q + u <= len); rep(q, u) counts the copies of U
// from q on. Sites of a class:
//   deletion, u = 1, base class B, n copies (1..6+): run start, B(b[q]) = B,
//       rep = n;
//   deletion, u >= 2, r repeats (0..5+): run start, rep = r + 1, and for
//       r = 0 no microhomology;
//   deletion with microhomology h, u >= 2: run start, rep = 1, and the first
//       h bases of U equal the h bases after it (h < u);
//   insertion with n copies (1..5+): run start, rep = n (u = 1: B(b[q]) = B);
//       the inserted unit is U;
//   insertion next to 0 copies: an inserted unit s (u = 1: one of the two
//       bases of class B, payload bit; u >= 2: any u-mer, payload = its 2u
//       bits) different from b[q..q+u) and from b[q-u..q).
// B is the pyrimidine of the base's pair (C for C/G, T for T/A). Classes are
// numbered in COSMIC order (PSX_ID_CLASSES in psx.h).
#pragma once
#include <stdint.h>

#include "psx_rng.h"

namespace psxid {

enum : uint32_t { DEL1 = 0, INS1 = 12, DELR = 24, INSR = 48, MH = 72, N_CLASSES = 83 };
// a class whose rejection search would need more tries than this on average has its sites listed
constexpr uint32_t DENSE_TRIES = 1024u;

PSX_HD uint32_t mh_base(uint32_t u) { return u == 2 ? 72u : (u == 3 ? 73u : (u == 4 ? 75u : 78u)); }
PSX_HD uint32_t mh_max(uint32_t u) { return u == 2 ? 1u : (u == 3 ? 2u : (u == 4 ? 3u : 5u)); }

struct cls_t {
    uint32_t del, u, b, n, h, zero;
};

PSX_HD cls_t decode(uint32_t c) {
    cls_t k{0, 0, 0, 0, 0, 0};
    if (c < INS1) { k.del = 1; k.u = 1; k.b = c / 6; k.n = c % 6 + 1; }
    else if (c < DELR) { k.u = 1; k.b = (c - INS1) / 6; k.n = (c - INS1) % 6; k.zero = k.n == 0; }
    else if (c < INSR) { k.del = 1; k.u = 2 + (c - DELR) / 6; k.n = (c - DELR) % 6 + 1; }
    else if (c < MH) { k.u = 2 + (c - INSR) / 6; k.n = (c - INSR) % 6; k.zero = k.n == 0; }
    else {
        k.del = 1; k.n = 1;
        k.u = c >= 78u ? 5u : (c >= 75u ? 4u : (c >= 73u ? 3u : 2u));
        k.h = c - mh_base(k.u) + 1;
    }
    return k;
}

// payloads of a class: 2 (1 bp insertions next to 0 copies), 4^u (longer ones), 1 otherwise
PSX_HD uint32_t payloads(uint32_t c) {
    const cls_t k = decode(c);
    return k.zero ? (k.u == 1 ? 2u : 1u << (2 * k.u)) : 1u;
}

PSX_HD uint32_t pyr(uint32_t base) { return (base == 1u || base == 2u) ? 0u : 1u; }

// the inserted unit of a payload of a class next to 0 copies
PSX_HD uint32_t zero_unit(uint32_t u, uint32_t b, uint32_t payload) {
    if (u == 1) return b == 0 ? ((payload & 1u) ? 2u : 1u) : ((payload & 1u) ? 0u : 3u);
    return payload;
}

// The class-zero insertion classes, in the order of the counts of sites_at.
PSX_HD uint32_t zero_class(uint32_t i) { return i == 0 ? INS1 : (i == 1 ? INS1 + 6 : INSR + 6 * (i - 2)); }

// A chromosome's bases through an accessor: BASE(i) for 0 <= i < len.
template <class BASE>
struct view {
    BASE base;
    uint32_t len;

    PSX_HD uint32_t unit(uint32_t q, uint32_t u) const {
        uint32_t s = 0;
        for (uint32_t i = 0; i < u; ++i) s |= base(q + i) << (2 * i);
        return s;
    }
    PSX_HD uint32_t rep(uint32_t q, uint32_t u, uint32_t U, uint32_t cap) const {
        uint32_t n = 1;
        while (n < cap && (uint64_t)q + (uint64_t)(n + 1) * u <= len && unit(q + n * u, u) == U) ++n;
        return n;
    }
    PSX_HD uint32_t mh(uint32_t q, uint32_t u) const {
        uint32_t h = 0;
        while (h + 1 < u && (uint64_t)q + u + h < len && base(q + h) == base(q + u + h)) ++h;
        return h;
    }

    // whether (q, payload) is a site of class c
    PSX_HD bool is_site(uint32_t q, uint32_t payload, uint32_t c) const {
        const cls_t k = decode(c);
        const uint32_t u = k.u;
        if (q < u || (uint64_t)q + u > len) return false;
        const uint32_t cur = unit(q, u), prev = unit(q - u, u);
        if (k.zero) {
            if (payload >= payloads(c)) return false;
            const uint32_t s = zero_unit(u, k.b, payload);
            return s != cur && s != prev;
        }
        if (cur == prev) return false;  // not a run start
        if (u == 1 && pyr(cur) != k.b) return false;
        if (!k.del) return rep(q, u, cur, 5) == k.n;
        const uint32_t n = rep(q, u, cur, 6);
        if (u == 1) return n == k.n;
        if (c >= MH) {
            const uint32_t h = n == 1 ? mh(q, u) : 0u;
            return h >= 1 && (h < mh_max(u) ? h : mh_max(u)) == k.h;
        }
        if (n != k.n) return false;
        return n >= 2 || mh(q, u) == 0;
    }

    // The classes with copies that q is a site of (up to 10, written to cls,
    // their number returned) and the payload counts of the classes next to 0
    // copies, added to z[6] (u = 1 of class C, u = 1 of class T, u = 2..5).
    PSX_HD uint32_t sites_at(uint32_t q, uint32_t* cls, uint32_t* z) const {
        uint32_t n = 0;
        for (uint32_t u = 1; u <= 5; ++u) {
            if (q < u || (uint64_t)q + u > len) continue;
            const uint32_t cur = unit(q, u), prev = unit(q - u, u);
            if (u == 1) {
                // payload bases of class C are C, G; of class T are T, A
                z[0] += (cur != 1u && prev != 1u) + (cur != 2u && prev != 2u);
                z[1] += (cur != 3u && prev != 3u) + (cur != 0u && prev != 0u);
            } else {
                z[u] += (1u << (2 * u)) - (cur == prev ? 1u : 2u);
            }
            if (cur == prev) continue;
            const uint32_t n6 = rep(q, u, cur, 6), n5 = n6 < 5 ? n6 : 5;
            if (u == 1) {
                const uint32_t b = pyr(cur);
                cls[n++] = DEL1 + b * 6 + n6 - 1;
                cls[n++] = INS1 + b * 6 + n5;
            } else {
                if (n6 >= 2) {
                    cls[n++] = DELR + (u - 2) * 6 + n6 - 1;
                } else {
                    const uint32_t h = mh(q, u);
                    cls[n++] = h ? mh_base(u) + (h < mh_max(u) ? h : mh_max(u)) - 1 : DELR + (u - 2) * 6;
                }
                cls[n++] = INSR + (u - 2) * 6 + n5;
            }
        }
        return n;
    }
};

}  // namespace psxid
