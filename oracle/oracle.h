/*
 * CPU oracle: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the sampling model of mutation placement and
 * count-only sequencing. Only tests/, the smoke check and the CPU-baseline
 * leg of bench.py may load it, and only to check results against: the product library
 * never calls it. Structurally it is the naive form (one read at a time,
 * binary search over the whole mutation table, sequential placement), written
 * independently of the HIP kernels; it shares with them only the plain-data
 * structs of include/psx.h and the published Philox4x32-10 algorithm.
 *
 * Reference semantics followed (ProCESS sources; the RACES library that holds
 * the arithmetic is not available, so the sampling model is this project's
 * own and is pinned by closed forms, invariants and brute force in tests/):
 *   place_mutations      src/mutation_engine.cpp:1085-1113,
 *                        src/mutation_engine.hpp:148-206
 *   simulate_seq         src/seq_simulation.cpp:354-451,517-601,
 *                        src/sequencing.cpp:193-209
 *   result table         src/seq_simulation.cpp:52-181
 */
#ifndef PSX_ORACLE_H
#define PSX_ORACLE_H

#include "../include/psx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_forest {
    uint32_t n_nodes, n_leaves, n_samples;
    int64_t* cell_id;      /* by DFS index */
    uint32_t* parent;      /* DFS index of the parent, PSX_NO_NODE for roots */
    uint32_t* species;
    double* birth;
    uint32_t* leaf_lo;
    uint32_t* leaf_hi;
    uint32_t* sample_begin;  /* n_samples + 1 */
    uint32_t* sample_leaves; /* leaf ranks, ascending within each sample */
} orc_forest;

const char* orc_last_error(void);
void orc_free(void* p);
/* OpenMP threads used by orc_simulate_seq (0 = runtime default); returns the
 * number that will be used. */
int orc_set_threads(int n);

uint64_t orc_genome_words(uint32_t n_chr, const uint32_t* len);
int orc_genome_synth(uint32_t n_chr, const uint32_t* len, uint64_t seed, uint32_t* packed);

int orc_forest_flatten(const psx_forest_desc* d, orc_forest** out);
void orc_forest_free(orc_forest* f);

int orc_place_mutations(uint32_t n_chr, const uint32_t* len, const uint32_t* packed,
                        const orc_forest* f, const psx_engine_params* p, uint64_t seed,
                        psx_mutation** out, uint64_t* n_out);

/* Copy-number drivers (psx_engine_params.cna) placed on the edge into every
 * node where their species first appears, in DFS order of the node, ties in
 * driver order (src/mutation_engine.cpp:1185-1210 for the driver order), and
 * the passenger CNAs of every non-root edge after that node's drivers. */
int orc_place_cna(uint32_t n_chr, const uint32_t* len, const orc_forest* f, const psx_engine_params* p,
                  uint64_t seed, psx_cna_event** out, uint32_t* n_out);

/* SBS context index of a reference at a sampling (0 counts as 1): sites per
 * pyrimidine-centred context (5' << 2 | 3', + 16 for a T centre), the kept
 * sites (every sampling-th in genome order, from the first) and their list
 * (32-bit global positions, ascending per context, contexts in order; free
 * with orc_free). */
int orc_sbs_index(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, uint32_t sampling,
                  uint64_t* counts, uint64_t* kept, uint32_t** list);
/* Indel (ID83) classes of a reference: sites per class, the classes whose
 * sites are listed (sparse: rejection would need more than 1024 tries on
 * average), list ranges off[84] and the lists (global position << 10 |
 * payload, ascending per class; free with orc_free). */
int orc_id_index(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, uint64_t* counts,
                 uint8_t* listed, uint64_t* off, uint64_t** list);
/* 1 when (chromosome position q, payload) is a site of ID class cls. */
int orc_id_site(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, uint32_t chr, uint32_t q,
                uint32_t payload, uint32_t cls);

/* Sorts an arbitrary table into SID order and marks sid_head. */
int orc_sort_table(psx_mutation* m, uint64_t n);

int orc_simulate_seq(uint32_t n_chr, const uint32_t* len, const orc_forest* f,
                     const psx_mutation* m, uint64_t n_mut, uint32_t ploidy,
                     const psx_seq_params* p, uint64_t seed, int32_t* alt, int32_t* depth,
                     uint8_t* present, uint64_t* n_reads);
/* The same with copy-number events (DFS order): reads come from allele
 * lineages in proportion to their copy numbers (the rejection model of
 * include/psx.h); *n_reads counts the reads of accepted templates. */
int orc_simulate_seq_cna(uint32_t n_chr, const uint32_t* len, const orc_forest* f,
                         const psx_mutation* m, uint64_t n_mut, uint32_t ploidy,
                         const psx_cna_event* ev, uint32_t n_ev,
                         const psx_seq_params* p, uint64_t seed, int32_t* alt, int32_t* depth,
                         uint8_t* present, uint64_t* n_reads);

/* simulate_seq in read-materialising form (psx_simulate_seq_reads): every
 * read into rows (capacity reads of 2 + ceil(L / 16) words each, in output
 * order: output sample, chromosome, stratum, template, mate; on a chromosome
 * with copy-number events only the accepted templates have reads, in template
 * order). */
int orc_simulate_reads(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, const orc_forest* f,
                       const psx_mutation* m, uint64_t n_mut, uint32_t ploidy, const psx_seq_params* p,
                       uint64_t seed, uint32_t* rows, uint64_t capacity, uint64_t* n_reads);
int orc_simulate_reads_cna(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, const orc_forest* f,
                           const psx_mutation* m, uint64_t n_mut, uint32_t ploidy, const psx_cna_event* ev,
                           uint32_t n_ev, const psx_seq_params* p, uint64_t seed, uint32_t* rows, uint64_t capacity,
                           uint64_t* n_reads);

/* Every read of simulate_seq (no copy-number events) handed to fn instead of
 * being counted: output sample, chromosome, start, source leaf (n_leaves for
 * the normal genome), allele, mate and the read's errors (offsets and
 * substitution indices). fn runs on the OpenMP threads of the walk. */
typedef void (*orc_read_fn)(void* user, uint32_t so, uint32_t chr, uint32_t x, uint32_t leaf, uint32_t allele,
                            uint32_t mate, const uint32_t* off, const uint32_t* sub, uint32_t n_err);
int orc_seq_walk(uint32_t n_chr, const uint32_t* len, const orc_forest* f, const psx_mutation* m, uint64_t n_mut,
                 uint32_t ploidy, const psx_seq_params* p, uint64_t seed, orc_read_fn fn, void* user,
                 uint64_t* n_reads);

/* The reference-shaped CPU restatement (refshape.cpp): the reads of
 * orc_seq_walk each built as base and quality strings from the reference,
 * their sources' mutations found in ordered maps and the statistics kept in
 * ordered maps, as RACES' ReadSimulator does; alt and depth [S][M] as
 * orc_simulate_seq. threads: OpenMP threads (1: the single-threaded shape). */
int orc_refshape_seq(uint32_t n_chr, const uint32_t* len, const uint32_t* packed, const orc_forest* f,
                     const psx_mutation* m, uint64_t n_mut, uint32_t ploidy, const psx_seq_params* p, uint64_t seed,
                     int threads, int32_t* alt, int32_t* depth, uint64_t* n_reads);

int orc_seq_table(const psx_mutation* recs, uint64_t n_mutations, uint32_t n_out_samples,
                  const int32_t* alt, const int32_t* depth, const uint8_t* present,
                  int32_t include_non_sequenced, uint64_t* n_rows, uint64_t* row_rec,
                  int32_t* occ, int32_t* cov, double* vaf);

/* Exposed for the known-answer tests. */
void orc_philox(const uint32_t ctr[4], uint64_t key, uint32_t out[4]);
/* Integer cumulative table of a pmf, as used by every inversion draw. */
int orc_cdf_table(const double* pmf, uint32_t n, uint64_t* T);
/* Inversion tables (value = k0 + index): Poisson(lambda), Binomial(t, p). */
int orc_poisson_table(double lambda, uint32_t* k0, uint32_t* n, uint64_t* T, uint32_t cap);
int orc_binomial_table(uint32_t t, double p, uint32_t* k0, uint32_t* n, uint64_t* T,
                       uint32_t cap);

#ifdef __cplusplus
}
#endif
#endif
