// The reference-shaped CPU restatement of count-only sequencing (test and
// baseline infrastructure, like the rest of oracle/): the same reads as
// orc_simulate_seq (orc_seq_walk hands them over), each processed the way
// RACES' ReadSimulator does per read (ProCESS src/seq_simulation.cpp:371-372
// calls it; RACES itself is not in the reference tree): the read's bases and
// quality string are built from the reference, the source's mutations in the
// read's range are found in an ordered map and applied, the sequencer's errors
// are applied to the string, and the per-mutation statistics (coverage and
// occurrences) are kept in ordered maps keyed by mutation. The resulting alt
// and depth equal orc_simulate_seq's; only the work per read differs.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

extern "C" {
#include "oracle.h"
}

namespace {

const char BASES[4] = {'A', 'C', 'G', 'T'};

uint32_t base_index(char b) { return b == 'A' ? 0u : b == 'C' ? 1u : b == 'G' ? 2u : 3u; }

struct thread_stats {
    std::map<uint64_t, uint32_t> coverage;     // (output sample, record) -> reads covering it
    std::map<uint64_t, uint32_t> occurrences;  // (output sample, SID head) -> reads showing it
    std::string bases, quals;
};

struct refshape {
    const uint32_t* packed = nullptr;
    std::vector<uint64_t> chr_woff;
    const psx_mutation* m = nullptr;
    uint32_t L = 0, n_leaves = 0;
    int preneo = 0;
    // (chr, position) -> records at that position, in table order
    std::map<uint64_t, std::vector<uint32_t>> by_pos;
    std::vector<thread_stats> stats;

    bool carried(const psx_mutation& r, uint32_t leaf, uint32_t allele) const {
        if (r.allele != allele) return false;
        if (leaf == n_leaves) return r.nature == PSX_GERMINAL || (r.nature == PSX_PRENEOPLASTIC && preneo);
        return r.leaf_lo <= leaf && leaf < r.leaf_hi;
    }

    void read(uint32_t so, uint32_t c, uint32_t x, uint32_t leaf, uint32_t allele, const uint32_t* off,
              const uint32_t* sub, uint32_t ne) {
        thread_stats& T = stats[(size_t)omp_get_thread_num()];
        std::string& b = T.bases;
        b.resize(L);
        for (uint32_t i = 0; i < L; ++i) {
            const uint64_t p = (uint64_t)x + i;
            b[i] = BASES[(packed[chr_woff[c] + p / 16] >> (2 * (p % 16))) & 3u];
        }
        T.quals.assign(L, 'I');  // constant qualities (the basic sequencer's default)
        const uint64_t k0 = ((uint64_t)c << 32) | x, k1 = ((uint64_t)c << 32) | ((uint64_t)x + L);
        const auto first = by_pos.lower_bound(k0);
        // the source genome's bases: the first carried SNV at each position
        for (auto it = first; it != by_pos.end() && it->first < k1; ++it)
            for (uint32_t q : it->second) {
                const psx_mutation& r = m[q];
                if (r.kind == PSX_SNV && carried(r, leaf, allele)) {
                    b[r.pos - x] = BASES[r.alt];
                    break;
                }
            }
        // the sequencer's errors
        for (uint32_t e = 0; e < ne; ++e) {
            const uint32_t tb = base_index(b[off[e]]);
            b[off[e]] = BASES[sub[e] < tb ? sub[e] : sub[e] + 1];
        }
        // statistics of every mutation the read covers
        for (auto it = first; it != by_pos.end() && it->first < k1; ++it) {
            const std::vector<uint32_t>& g = it->second;
            for (size_t k = 0; k < g.size(); ++k) {
                const uint32_t q = g[k];
                const psx_mutation& r = m[q];
                const uint64_t key = ((uint64_t)so << 40) | q;
                ++T.coverage[key];
                if (!r.sid_head) continue;
                bool seen;
                if (r.kind == PSX_SNV) {
                    seen = b[r.pos - x] == BASES[r.alt];
                } else {
                    seen = false;
                    for (size_t k2 = k; k2 < g.size() && (k2 == k || !m[g[k2]].sid_head); ++k2)
                        seen = seen || carried(m[g[k2]], leaf, allele);
                }
                if (seen) ++T.occurrences[key];
            }
        }
    }
};

void on_read(void* user, uint32_t so, uint32_t c, uint32_t x, uint32_t leaf, uint32_t allele, uint32_t mate,
             const uint32_t* off, const uint32_t* sub, uint32_t ne) {
    (void)mate;
    static_cast<refshape*>(user)->read(so, c, x, leaf, allele, off, sub, ne);
}

}  // namespace

extern "C" int orc_refshape_seq(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, const orc_forest* f,
                                const psx_mutation* m, uint64_t n_mut, uint32_t ploidy, const psx_seq_params* p,
                                uint64_t seed, int threads, int32_t* alt, int32_t* depth, uint64_t* n_reads) {
    refshape R;
    R.packed = packed;
    R.m = m;
    R.L = p->read_size;
    R.n_leaves = f->n_leaves;
    R.preneo = p->preneoplastic_in_normal != 0;
    R.chr_woff.resize(n_chr);
    uint64_t w = 0;
    for (uint32_t c = 0; c < n_chr; ++c) {
        R.chr_woff[c] = w;
        w += (len[c] + 15u) / 16u;
    }
    for (uint64_t q = 0; q < n_mut; ++q) R.by_pos[((uint64_t)m[q].chr << 32) | m[q].pos].push_back((uint32_t)q);
    const int prev = orc_set_threads(0);
    orc_set_threads(threads > 0 ? threads : prev);
    R.stats.resize((size_t)std::max(omp_get_max_threads(), 1));
    const int rc = orc_seq_walk(n_chr, len, f, m, n_mut, ploidy, p, seed, on_read, &R, n_reads);
    orc_set_threads(prev);
    if (rc) return rc;
    const uint32_t n_out = (p->samples ? p->n_samples : f->n_samples) + (p->with_normal_sample ? 1u : 0u);
    std::memset(alt, 0, sizeof(int32_t) * n_out * n_mut);
    std::memset(depth, 0, sizeof(int32_t) * n_out * n_mut);
    for (const thread_stats& T : R.stats) {
        for (const auto& kv : T.coverage) depth[(kv.first >> 40) * n_mut + (kv.first & 0xFFFFFFFFFFull)] += (int32_t)kv.second;
        for (const auto& kv : T.occurrences) alt[(kv.first >> 40) * n_mut + (kv.first & 0xFFFFFFFFFFull)] += (int32_t)kv.second;
    }
    return 0;
}
