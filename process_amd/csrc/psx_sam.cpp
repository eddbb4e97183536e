// SAM files of simulate_seq(write_SAM = TRUE) (src/seq_simulation.cpp:536-564,
// vignettes/sequencing.Rmd:137-180, 272-309), written from the device reads
// of psx_simulate_seq_reads without a host copy of the whole read set: one
// chromosome's reads at a time are made on the GPU, copied back through two
// pinned chunks (the next chunk's copy runs while the current one is
// formatted) and formatted into SAM lines by several host threads.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "psx_internal.h"
#include "psx_seq_host.h"

namespace {

constexpr uint64_t CHUNK_READS = 1u << 20;
constexpr uint32_t MAX_THREADS = 16;

bool file_exists(const std::string& p) {
    struct stat st;
    return ::stat(p.c_str(), &st) == 0;
}

struct sam_file {
    FILE* fh = nullptr;
    ~sam_file() {
        if (fh) std::fclose(fh);
    }
};

struct line_ctx {
    const char* chr_name;
    std::string qname_prefix;
    uint32_t L, rw, rpt;
    int random_q;
    char q_const;
    uint32_t q0;
    uint64_t seed;
};

// One read's SAM line; rows holds its mate's row too (paired ends).
void format_read(const line_ctx& x, const uint32_t* row, const uint32_t* mate_row, uint64_t tmpl, uint64_t full_row,
                 const char* rg, std::string& out) {
    static const char BASES[4] = {'A', 'C', 'G', 'T'};
    const uint32_t start = row[0];
    char num[32];
    out += x.qname_prefix;
    std::snprintf(num, sizeof num, "%llu", (unsigned long long)tmpl);
    out += num;
    if (x.rpt == 2) {
        const uint32_t mate = (row[1] >> 30) & 1u;
        const uint32_t other = mate_row[0];
        const uint64_t lo = std::min(start, other), hi = (uint64_t)std::max(start, other) + x.L;
        const long long tlen = mate == 0 ? (long long)(hi - lo) : -(long long)(hi - lo);
        std::snprintf(num, sizeof num, "\t%u\t", mate == 0 ? 99u : 147u);
        out += num;
        out += x.chr_name;
        std::snprintf(num, sizeof num, "\t%u\t60\t", start + 1);
        out += num;
        std::snprintf(num, sizeof num, "%uM\t=\t", x.L);
        out += num;
        std::snprintf(num, sizeof num, "%u\t%lld\t", other + 1, tlen);
        out += num;
    } else {
        out += "\t0\t";
        out += x.chr_name;
        std::snprintf(num, sizeof num, "\t%u\t60\t%uM\t*\t0\t0\t", start + 1, x.L);
        out += num;
    }
    const size_t s0 = out.size();
    out.resize(s0 + x.L);
    for (uint32_t i = 0; i < x.L; ++i) out[s0 + i] = BASES[(row[2 + i / 16] >> (2u * (i % 16))) & 3u];
    out += '\t';
    const size_t q0 = out.size();
    out.resize(q0 + x.L);
    if (!x.random_q) {
        std::memset(&out[q0], x.q_const, x.L);
    } else {
        for (uint32_t j = 0; j * 16 < x.L; ++j) {
            const psx_u4 w = psx_draw(x.seed, (uint32_t)full_row, (uint32_t)(full_row >> 32), j, PSX_P_QUALITY << 24);
            const uint32_t words[4] = {w.x, w.y, w.z, w.w};
            for (uint32_t b = 0; b < 16 && j * 16 + b < x.L; ++b) {
                const uint32_t byte = (words[b / 4] >> (8 * (b % 4))) & 0xFFu;
                int q = (int)x.q0 - 5 + (int)(byte % 11u);
                q = std::min(41, std::max(2, q));
                out[q0 + j * 16 + b] = (char)(33 + q);
            }
        }
    }
    out += "\tRG:Z:";
    out += rg;
    out += '\n';
}

}  // namespace

extern "C" psx_status psx_write_sam(psx_ctx* ctx, const psx_genome* g, const psx_forest* f, const psx_mutations* m,
                                    const psx_seq_params* p, uint64_t seed, const psx_sam_params* sam,
                                    uint64_t* n_reads) {
    if (!ctx || !g || !f || !m || !p || !sam || !sam->output_dir || !sam->sample_names || !sam->chr_names)
        return PSX_EINVAL;
    void* staging = nullptr;
    uint32_t* d_reads = nullptr;
    psx_status st = psx_guard(ctx, [&] {
        PSX_HIP(hipSetDevice(ctx->device));
        const uint32_t L = p->read_size;
        const bool pe = p->insert_size_mean > 0;
        const uint32_t rpt = pe ? 2u : 1u;
        const uint32_t S = psx_seq_num_samples(f, p);
        const uint32_t rw = psx_read_words(p);
        if (L < 1) psx_throw(PSX_EINVAL, "invalid read size");
        // selected chromosomes, in index order, with room for a read
        std::vector<uint8_t> sel(g->n_chr, p->n_chromosomes == 0 ? 1 : 0);
        for (uint32_t k = 0; k < p->n_chromosomes; ++k) {
            if (p->chromosomes[k] >= g->n_chr) psx_throw(PSX_EINVAL, "unknown chromosome %u", p->chromosomes[k]);
            sel[p->chromosomes[k]] = 1;
        }
        std::vector<uint32_t> chrs;
        for (uint32_t c = 0; c < g->n_chr; ++c)
            if (sel[c] && g->len[c] >= L) chrs.push_back(c);
        // reads of every (output sample, chromosome) segment (kept templates with copy-number events), the
        // first read of each in the order of one call over every selected chromosome, and the largest call
        const std::vector<uint64_t> seg = psxh::segment_reads(ctx, psxh::job_in{g, f, m, p, seed, nullptr, nullptr,
                                                                                nullptr});
        std::vector<uint64_t> seg_first(seg.size(), 0);
        uint64_t max_reads = 0;
        {
            uint64_t at = 0;
            for (uint32_t so = 0; so < S; ++so)
                for (uint32_t c : chrs) {
                    seg_first[(size_t)so * g->n_chr + c] = at;
                    at += seg[(size_t)so * g->n_chr + c];
                }
            for (uint32_t c : chrs) {
                uint64_t n = 0;
                for (uint32_t so = 0; so < S; ++so) n += seg[(size_t)so * g->n_chr + c];
                max_reads = std::max(max_reads, n);
            }
        }
        // file names first: nothing is written when one is taken (create mode)
        const std::string dir = sam->output_dir;
        std::vector<std::string> paths;
        for (uint32_t c : chrs) {
            const std::string stem = dir + "/" + (sam->filename_prefix ? sam->filename_prefix : "") + sam->chr_names[c];
            std::string path = stem + ".sam";
            if (file_exists(path)) {
                if (!sam->update) psx_throw(PSX_EINVAL, "the SAM file \"%s\" already exists", path.c_str());
                for (uint64_t n = 1;; ++n) {
                    path = stem + "_" + std::to_string(n) + ".sam";
                    if (!file_exists(path)) break;
                }
            }
            paths.push_back(path);
        }
        const double e = p->error_rate;
        line_ctx x{};
        x.qname_prefix = sam->template_name_prefix ? sam->template_name_prefix : "";
        x.L = L; x.rw = rw; x.rpt = rpt; x.random_q = sam->random_quality; x.seed = seed;
        x.q0 = e <= 0.0 ? 40u : (uint32_t)std::min(40.0, std::max(0.0, std::round(-10.0 * std::log10(e))));
        x.q_const = (char)(33 + x.q0);
        const uint64_t chunk = std::min<uint64_t>(CHUNK_READS, std::max<uint64_t>(max_reads, 1));
        const size_t chunk_bytes = (size_t)chunk * rw * 4u;
        PSX_HIP(hipHostMalloc(&staging, 2 * chunk_bytes, hipHostMallocDefault));
        if (max_reads) PSX_HIP(hipMalloc((void**)&d_reads, (size_t)max_reads * rw * 4u));
        const uint32_t n_threads = std::max(1u, std::min(MAX_THREADS, std::thread::hardware_concurrency()));
        uint64_t total = 0;
        for (size_t ci = 0; ci < chrs.size(); ++ci) {
            psx_check_cancel(ctx);
            const uint32_t c = chrs[ci];
            psx_seq_params pc = *p;
            pc.n_chromosomes = 1;
            pc.chromosomes = &c;
            uint64_t n = 0;
            const psx_status rs = psx_simulate_seq_reads(ctx, g, f, m, &pc, seed, d_reads, max_reads, &n);
            if (rs != PSX_OK) psx_throw(rs, "%s", ctx->err.c_str());
            // this call's reads: the chromosome's segments, output sample after output sample
            std::vector<uint64_t> call_first(S + 1, 0);
            for (uint32_t so = 0; so < S; ++so) call_first[so + 1] = call_first[so] + seg[(size_t)so * g->n_chr + c];
            if (n != call_first[S]) psx_throw(PSX_EHIP, "chromosome %u: %llu reads, expected %llu", c,
                                              (unsigned long long)n, (unsigned long long)call_first[S]);
            sam_file out;
            out.fh = std::fopen(paths[ci].c_str(), "w");
            if (!out.fh) psx_throw(PSX_EINVAL, "cannot open \"%s\" for writing", paths[ci].c_str());
            std::string head = "@HD\tVN:1.6\tSO:unsorted\n@SQ\tSN:";
            head += sam->chr_names[c];
            head += "\tLN:" + std::to_string(g->len[c]) + "\n";
            for (uint32_t so = 0; so < S; ++so) {
                head += "@RG\tID:";
                head += sam->sample_names[so];
                head += "\tSM:";
                head += sam->sample_names[so];
                head += "\n";
            }
            std::fwrite(head.data(), 1, head.size(), out.fh);
            x.chr_name = sam->chr_names[c];
            // chunks of whole templates; the next chunk's copy overlaps the current chunk's formatting
            const uint64_t chunk_t = std::max<uint64_t>(rpt, chunk / rpt * rpt);
            const uint64_t n_chunks = (n + chunk_t - 1) / chunk_t;
            auto issue = [&](uint64_t k) {
                const uint64_t r0 = k * chunk_t, cnt = std::min(chunk_t, n - r0);
                uint8_t* dst = (uint8_t*)staging + (k & 1) * chunk_bytes;
                PSX_HIP(hipMemcpyAsync(dst, d_reads + r0 * rw, cnt * rw * 4u, hipMemcpyDeviceToHost, ctx->stream));
                PSX_HIP(hipEventRecord((k & 1) ? ctx->ev1 : ctx->ev0, ctx->stream));
            };
            if (n_chunks) issue(0);
            std::vector<std::string> parts(n_threads);
            for (uint64_t k = 0; k < n_chunks; ++k) {
                psx_check_cancel(ctx);
                PSX_HIP(hipEventSynchronize((k & 1) ? ctx->ev1 : ctx->ev0));
                if (k + 1 < n_chunks) issue(k + 1);
                const uint64_t r0 = k * chunk_t, cnt = std::min(chunk_t, n - r0);
                const uint32_t* rows = (const uint32_t*)((const uint8_t*)staging + (k & 1) * chunk_bytes);
                const uint64_t per = (cnt / rpt + n_threads - 1) / n_threads * rpt;
                std::vector<std::thread> th;
                for (uint32_t t = 0; t < n_threads; ++t) {
                    th.emplace_back([&, t] {
                        std::string& s = parts[t];
                        s.clear();
                        const uint64_t a = std::min<uint64_t>(cnt, t * per), b = std::min<uint64_t>(cnt, a + per);
                        s.reserve((size_t)(b - a) * (2 * L + 96));
                        uint32_t so = 0;
                        for (uint64_t i = a; i < b; ++i) {
                            const uint64_t r = r0 + i;  // read of this chromosome's call
                            while (r >= call_first[so + 1]) ++so;
                            const uint64_t kin = r - call_first[so];
                            // the read's index in a call over every selected chromosome, and its template's
                            const uint64_t full_row = seg_first[(size_t)so * g->n_chr + c] + kin;
                            const uint64_t tmpl = full_row / rpt;
                            const uint32_t* row = rows + i * rw;
                            const uint32_t* mrow = rpt == 2 ? rows + (i ^ 1u) * rw : row;
                            format_read(x, row, mrow, tmpl, full_row, sam->sample_names[so], s);
                        }
                    });
                }
                for (auto& t : th) t.join();
                for (const std::string& s : parts)
                    if (!s.empty() && std::fwrite(s.data(), 1, s.size(), out.fh) != s.size())
                        psx_throw(PSX_EINVAL, "writing \"%s\" failed", paths[ci].c_str());
            }
            if (std::fclose(out.fh) != 0) {
                out.fh = nullptr;
                psx_throw(PSX_EINVAL, "closing \"%s\" failed", paths[ci].c_str());
            }
            out.fh = nullptr;
            total += n;
        }
        if (n_reads) *n_reads = total;
    });
    if (st != PSX_OK) (void)hipStreamSynchronize(ctx->stream);  // a chunk's copy may still be in flight
    if (d_reads) (void)hipFree(d_reads);
    if (staging) (void)hipHostFree(staging);
    return st;
}
