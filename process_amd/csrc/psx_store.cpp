// Binary save and load of a device mutation table: checkpoint and resume for
// the hot path's input (a placed table sequenced again by another process).
// The reference saves a PhylogeneticForest as a RACES archive
// (src/phylogenetic_forest.cpp:940-992); that format belongs to RACES, so this
// is the library's own. Its loader's checks and error texts follow the
// reference's (src/phylogenetic_forest.cpp:958-986, src/utility.cpp:141-161).
//
// Layout, little endian:
//   char     magic[8]            "PSXMUTBL"
//   uint32   version             PSX_TABLE_FORMAT_VERSION
//   uint32   ploidy, n_leaves, n_chr
//   uint64   n                   records
//   uint32   n_cna, reserved (0)
//   uint32   chr_len[n_chr]      chromosome lengths at placement
//   psx_mutation records[n]      SID order, sid_head set
//   psx_cna_event events[n_cna]  DFS order of their nodes
//   uint64   checksum            of every byte before it (a word-wise FNV-1a)
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "psx_internal.h"

namespace {

constexpr char MAGIC[8] = {'P', 'S', 'X', 'M', 'U', 'T', 'B', 'L'};

struct header {
    char magic[8];
    uint32_t version, ploidy, n_leaves, n_chr;
    uint64_t n;
    uint32_t n_cna, reserved;
};
static_assert(sizeof(header) == 40, "header layout");

// FNV-1a over 64-bit words (the tail zero-padded)
struct checksum {
    uint64_t h = 0xcbf29ce484222325ull;
    void add(const void* p, size_t bytes) {
        const uint8_t* b = (const uint8_t*)p;
        size_t i = 0;
        for (; i + 8 <= bytes; i += 8) {
            uint64_t w;
            std::memcpy(&w, b + i, 8);
            h = (h ^ w) * 0x100000001b3ull;
        }
        if (i < bytes) {
            uint64_t w = 0;
            std::memcpy(&w, b + i, bytes - i);
            h = (h ^ w) * 0x100000001b3ull;
        }
    }
};

struct file_closer {
    std::FILE* f;
    ~file_closer() {
        if (f) std::fclose(f);
    }
};

[[noreturn]] void wrong_format(const std::string& path) {
    psx_throw(PSX_EINVAL, "Wrong file format for the mutation table file \"%s\". Remove it and retry.", path.c_str());
}

}  // namespace

extern "C" psx_status psx_mutations_save(psx_ctx* ctx, const psx_mutations* m, const char* path) {
    if (!ctx || !m || !path) return PSX_EINVAL;
    return psx_guard(ctx, [&] {
        std::vector<psx_mutation> recs(m->n);
        if (m->n) {
            PSX_HIP(hipSetDevice(ctx->device));
            PSX_HIP(hipMemcpyAsync(recs.data(), m->recs, m->n * sizeof(psx_mutation), hipMemcpyDeviceToHost,
                                   ctx->stream));
            PSX_HIP(hipStreamSynchronize(ctx->stream));
        }
        header h{};
        std::memcpy(h.magic, MAGIC, 8);
        h.version = PSX_TABLE_FORMAT_VERSION;
        h.ploidy = m->ploidy;
        h.n_leaves = m->n_leaves;
        h.n_chr = (uint32_t)m->chr_len.size();
        h.n = m->n;
        h.n_cna = (uint32_t)m->cna.size();
        file_closer fc{std::fopen(path, "wb")};
        if (!fc.f) psx_throw(PSX_EINVAL, "The file \"%s\" cannot be written.", path);
        checksum cs;
        auto put = [&](const void* p, size_t bytes) {
            if (bytes && std::fwrite(p, 1, bytes, fc.f) != bytes)
                psx_throw(PSX_EINVAL, "Writing the file \"%s\" failed.", path);
            cs.add(p, bytes);
        };
        put(&h, sizeof h);
        put(m->chr_len.data(), m->chr_len.size() * sizeof(uint32_t));
        put(recs.data(), recs.size() * sizeof(psx_mutation));
        put(m->cna.data(), m->cna.size() * sizeof(psx_cna_event));
        const uint64_t sum = cs.h;
        put(&sum, sizeof sum);
        if (std::fclose(fc.f) != 0) {
            fc.f = nullptr;
            psx_throw(PSX_EINVAL, "Writing the file \"%s\" failed.", path);
        }
        fc.f = nullptr;
    });
}

extern "C" psx_status psx_mutations_load(psx_ctx* ctx, const char* path, psx_mutations** out) {
    if (!ctx || !path || !out) return PSX_EINVAL;
    psx_mutations* m = new psx_mutations();
    m->ctx = ctx;
    const psx_status st = psx_guard(ctx, [&] {
        const std::string p(path);
        std::error_code ec;
        if (!std::filesystem::exists(p, ec)) psx_throw(PSX_EINVAL, "The file \"%s\" does not exist.", path);
        if (!std::filesystem::is_regular_file(p, ec))
            psx_throw(PSX_EINVAL, "The file \"%s\" is not a regular file.", path);
        const uint64_t size = std::filesystem::file_size(p, ec);
        file_closer fc{std::fopen(path, "rb")};
        if (!fc.f) psx_throw(PSX_EINVAL, "The file \"%s\" cannot be read.", path);
        checksum cs;
        auto get = [&](void* q, size_t bytes) {
            if (bytes && std::fread(q, 1, bytes, fc.f) != bytes) wrong_format(p);
            cs.add(q, bytes);
        };
        header h{};
        if (size < sizeof h) wrong_format(p);
        get(&h, sizeof h);
        if (std::memcmp(h.magic, MAGIC, 8) != 0) wrong_format(p);
        if (h.version != PSX_TABLE_FORMAT_VERSION)
            psx_throw(PSX_EINVAL,
                      "The mutation table file \"%s\" is in the file format version %u, expected version %u. "
                      "Remove it and retry.",
                      path, h.version, (unsigned)PSX_TABLE_FORMAT_VERSION);
        if (h.ploidy < 1 || h.ploidy > 255 || h.n_chr > PSX_MAX_CHROMOSOMES || h.reserved != 0) wrong_format(p);
        const uint64_t want = sizeof h + 4ull * h.n_chr + (uint64_t)sizeof(psx_mutation) * h.n +
                              (uint64_t)sizeof(psx_cna_event) * h.n_cna + 8u;
        if (h.n > (size / sizeof(psx_mutation)) || want != size) wrong_format(p);
        std::vector<uint32_t> chr_len(h.n_chr);
        get(chr_len.data(), chr_len.size() * 4u);
        std::vector<psx_mutation> recs(h.n);
        get(recs.data(), recs.size() * sizeof(psx_mutation));
        std::vector<psx_cna_event> ev(h.n_cna);
        get(ev.data(), ev.size() * sizeof(psx_cna_event));
        const uint64_t expect = cs.h;
        uint64_t sum = 0;
        get(&sum, sizeof sum);
        if (sum != expect) wrong_format(p);
        for (const psx_mutation& r : recs) {
            if (r.chr >= h.n_chr || r.pos >= chr_len[r.chr] || r.allele >= h.ploidy || r.kind > PSX_INS ||
                r.nature > PSX_GERMINAL || r.ref > 3 || r.alt > 3 || r.len > 15 || r.leaf_lo > r.leaf_hi ||
                r.leaf_hi > h.n_leaves)
                wrong_format(p);
        }
        m->ploidy = h.ploidy;
        m->n_leaves = h.n_leaves;
        PSX_HIP(hipSetDevice(ctx->device));
        psx_mutations_build(ctx, m, recs, chr_len);
        psx_cna_check(ev, m->n_leaves, m->ploidy, m->chr_len);
        m->cna = std::move(ev);
    });
    if (st != PSX_OK) {
        delete m;
        return st;
    }
    *out = m;
    return PSX_OK;
}
