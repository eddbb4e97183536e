// Copy numbers of a cell's allele lineages from placed copy-number events,
// shared by the host (launch bounds) and the sequencing kernels.
//
// A leaf starts with one copy of each of its `ploidy` allele lineages; the
// events of its lineage apply in order (events are kept in DFS order of their
// node, so an ancestor's events come first): an amplification adds copies to
// one lineage, a deletion removes all of them, a whole-genome doubling
// doubles every lineage. The normal genome has no events.
#pragma once
#include <stdint.h>

#include "../../include/psx.h"
#include "psx_rng.h"

// Device form of a placed event.
struct psx_cna_dev {
    uint32_t lo, hi;      // leaf range
    uint32_t start, end;  // [start, end); the whole chromosome for a doubling
    uint32_t code;        // allele | kind << 8 | copies << 16
};

PSX_HD uint32_t psx_cna_code(const psx_cna_event& e) {
    return (uint32_t)e.allele | ((uint32_t)e.kind << 8) | ((uint32_t)e.copies << 16);
}

// One event applied to the copy numbers cn[0 .. ploidy) of a leaf it covers.
PSX_HD void psx_cna_apply(uint32_t* cn, uint32_t ploidy, uint32_t code) {
    const uint32_t a = code & 0xFFu, kind = (code >> 8) & 0xFFu, c = code >> 16;
    if (kind == PSX_CNA_WGD) {
        for (uint32_t q = 0; q < ploidy; ++q) cn[q] *= 2u;
    } else if (kind == PSX_CNA_AMP) {
        cn[a] += c;
    } else {
        cn[a] = 0u;
    }
}

// Copy numbers of `leaf` at base x from the chromosome's events; returns the
// total.
PSX_HD uint32_t psx_cna_copies(const psx_cna_dev* ev, uint32_t n, uint32_t leaf, uint32_t x, uint32_t ploidy,
                               uint32_t* cn) {
    for (uint32_t a = 0; a < ploidy; ++a) cn[a] = 1u;
    for (uint32_t i = 0; i < n; ++i)
        if (leaf >= ev[i].lo && leaf < ev[i].hi && x >= ev[i].start && x < ev[i].end)
            psx_cna_apply(cn, ploidy, ev[i].code);
    uint32_t total = 0;
    for (uint32_t a = 0; a < ploidy; ++a) total += cn[a];
    return total;
}
