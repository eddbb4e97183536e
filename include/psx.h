/*
 * psx.h: C ABI of the MI355X (gfx950) mutation-placement and sequencing
 * library. Plain pointers and sizes only; no exceptions cross this boundary.
 *
 * Every entry point returns a psx_status; on failure psx_last_error(ctx)
 * holds the message (the Rcpp shim rethrows it as an R error, which is how
 * the reference reports failures: src/seq_simulation.cpp:103,254-259,437-445,
 * src/sequencers.cpp:86-101).
 *
 * One psx_ctx owns one HIP stream and is used by one host thread at a time;
 * several contexts may run concurrently. Genomes, forests and mutation tables
 * are opaque device-resident handles that stay in HBM across calls.
 *
 * Reference interfaces replaced (paths relative to the ProCESS sources):
 *   psx_forest_upload      SamplesForest node table, src/forest.hpp:29-67,
 *                          src/samples_forest.hpp:31-82
 *   psx_place_mutations    MutationEngine::place_mutations,
 *                          src/mutation_engine.cpp:1085-1113 (RACES call at
 *                          :1097-1101), overloads src/mutation_engine.hpp:160-206
 *   psx_simulate_seq       the ReadSimulator call inside simulate_seq,
 *                          src/seq_simulation.cpp:354-429 (RACES call at
 *                          :371-372, :413-414, :423-424)
 *   psx_seq_table          get_result_dataframe / get_active_mutations /
 *                          add_sample_statistics, src/seq_simulation.cpp:52-181
 *   psx_insert_size_check  get_bin_dist, src/seq_simulation.cpp:431-451
 */
#ifndef PSX_H
#define PSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSX_ABI_VERSION 1

/* Reads whose template starts fall in one stratum of PSX_STRATUM reference
 * positions are drawn from one counter range; part of the sampling model. */
#define PSX_STRATUM 8192u
/* Sequencing errors are independent per base and drawn per template (both
 * mates); an error rate is accepted when a template holds more than
 * PSX_ERR_KCAP errors with probability below 2^-40. Part of the sampling
 * model. */
#define PSX_ERR_KCAP 128u
#define PSX_MAX_CHROMOSOMES 64u
#define PSX_NO_NODE 0xFFFFFFFFu
#define PSX_NO_CAUSE 0xFFFFu

typedef enum psx_status {
    PSX_OK = 0,
    PSX_EINVAL = 1,          /* invalid argument (message in psx_last_error) */
    PSX_ENOMEM_DEVICE = 2,   /* device allocation failed */
    PSX_EHIP = 3,            /* HIP runtime error */
    PSX_ECANCELLED = 4,      /* cancelled between work chunks */
    PSX_ENOMEM_HOST = 5,
    PSX_EUNSUPPORTED = 6
} psx_status;

/* Mutation kinds and natures (the reference's "classes" column:
 * src/phylogenetic_forest.cpp:342-344). */
enum { PSX_SNV = 0, PSX_DEL = 1, PSX_INS = 2 };
enum { PSX_PASSENGER = 0, PSX_DRIVER = 1, PSX_PRENEOPLASTIC = 2, PSX_GERMINAL = 3 };

typedef struct psx_ctx psx_ctx;
typedef struct psx_genome psx_genome;
typedef struct psx_forest psx_forest;
typedef struct psx_mutations psx_mutations;

/* ------------------------------------------------------------------ context */

int psx_abi_version(void);
psx_status psx_ctx_create(int device, psx_ctx** out);
void psx_ctx_destroy(psx_ctx* ctx);
const char* psx_last_error(const psx_ctx* ctx);
psx_status psx_ctx_synchronize(psx_ctx* ctx);
/* Milliseconds of the dominant kernel(s) of the last sequencing call, timed
 * with HIP events on the context stream. */
psx_status psx_ctx_last_kernel_ms(const psx_ctx* ctx, double* ms);
/* Free and total HBM of the context device, in bytes. */
psx_status psx_ctx_mem_info(psx_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes);
/* Frees the buffers a context keeps grown between calls (placement
 * temporaries, psx_simulate_seq's device tables, launch descriptors, pinned
 * staging) after waiting for its stream; later calls grow them again. No
 * counterpart in the reference: for long-lived processes (an R session) that
 * ran one large call and want the memory back. */
psx_status psx_ctx_trim(psx_ctx* ctx);

/* Device buffers on the context device for callers that have no allocator of
 * their own in this process (no reference counterpart; plumbing for the
 * device-pointer entry points). Copies are synchronous. */
psx_status psx_dev_alloc(psx_ctx* ctx, uint64_t bytes, void** out);
void psx_dev_free(psx_ctx* ctx, void* p);
psx_status psx_dev_upload(psx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
psx_status psx_dev_download(psx_ctx* ctx, void* dst, const void* src, uint64_t bytes);
psx_status psx_dev_fill_u8(psx_ctx* ctx, void* dst, uint8_t value, uint64_t bytes);
psx_status psx_dev_copy(psx_ctx* ctx, void* dst, const void* src, uint64_t bytes);

/* ------------------------------------------------------------------- genome */

/* Synthetic reference: base i of chromosome c is i.i.d. uniform over ACGT,
 * drawn from Philox(key = ref_seed). Stored 2-bit packed (A=0 C=1 G=2 T=3,
 * 16 bases per uint32, base i in bits 2*(i%16)), each chromosome starting on a
 * fresh word. Replaces the FASTA the reference downloads
 * (src/mutation_engine.cpp:75-112), which needs network access. */
psx_status psx_genome_synth(psx_ctx* ctx, uint32_t n_chr, const uint32_t* chr_len,
                            uint64_t ref_seed, psx_genome** out);
/* A caller-provided packed reference in the same layout. */
psx_status psx_genome_upload(psx_ctx* ctx, uint32_t n_chr, const uint32_t* chr_len,
                             const uint32_t* packed, psx_genome** out);
/* Words of the packed layout: sum over chromosomes of ceil(len / 16). */
uint64_t psx_genome_words(uint32_t n_chr, const uint32_t* chr_len);
psx_status psx_genome_download(psx_ctx* ctx, const psx_genome* g, uint32_t* packed,
                               uint64_t n_words);
void psx_genome_free(psx_genome* g);
/* The genome's SBS context index at a sampling (built on the GPU at the first
 * placement or call with that sampling, kept with the genome; replaces
 * build_contex_index, src/mutation_engine.cpp:213-261): counts[32] sites and
 * kept[32] indexed sites per pyrimidine-centred context (5' base << 2 | 3'
 * base, + 16 for a T centre), and, if list is not NULL, the indexed global
 * positions, ascending per context in context order (list_cap entries at
 * least: the sum of kept). build_ms: GPU time of the build. Any pointer but
 * ctx and g may be NULL. */
psx_status psx_genome_context_index(psx_ctx* ctx, psx_genome* g, uint32_t sampling, uint64_t* counts,
                                    uint64_t* kept, uint32_t* list, uint64_t list_cap, double* build_ms);

/* ------------------------------------------------------------------- forest */

/* One row per cell, as SamplesForest$get_nodes() (src/forest.hpp:29-67). */
typedef struct psx_forest_desc {
    uint32_t n_nodes;
    const int64_t* cell_id;
    const int64_t* parent;    /* ancestor cell id, -1 for a root */
    const uint32_t* species;  /* species (mutant x epistate) index */
    const int32_t* sample;    /* sample index of a sampled leaf, -1 otherwise */
    const double* birth_time;
    uint32_t n_samples;
} psx_forest_desc;

/* Flattens the forest into DFS pre-order (roots and children by ascending
 * cell id) and uploads it. A node's subtree holds the leaves with ranks
 * [leaf_lo, leaf_hi). */
psx_status psx_forest_upload(psx_ctx* ctx, const psx_forest_desc* desc, psx_forest** out);
psx_status psx_forest_info(const psx_forest* f, uint32_t* n_nodes, uint32_t* n_leaves,
                           uint32_t* n_samples);
/* Host copies of the flattened arrays (each n_nodes long, except
 * sample_leaves which is n_leaves long and sample_begin n_samples + 1). */
psx_status psx_forest_flat(const psx_forest* f, int64_t* dfs_cell_id, uint32_t* parent_dfs,
                           uint32_t* leaf_lo, uint32_t* leaf_hi, uint32_t* sample_begin,
                           uint32_t* sample_leaves);
void psx_forest_free(psx_forest* f);

/* ---------------------------------------------------------------- placement */

typedef struct psx_driver {
    uint32_t species;  /* the driver is placed on the edge into every node where this
                          species first appears on a lineage */
    uint32_t chr;
    uint32_t pos;      /* 0-based */
    uint8_t alt;       /* 0..3 */
    int8_t allele;     /* copy index, -1 = uniform draw */
    uint8_t pad[2];
} psx_driver;

typedef struct psx_germline {
    uint32_t chr;
    uint32_t pos;
    uint8_t alt;
    uint8_t allele;
    uint8_t pad[2];
} psx_germline;

/* Copy-number driver of a mutant: CNA("A"/"D", chr, pos, len, allele,
 * src_allele) and WholeGenomeDoubling of the reference (src/cna.hpp:28-94,
 * src/mutations.cpp:355-365, applied in the mutant's driver order,
 * src/mutation_engine.cpp:1185-1210). Copy numbers are kept per allele
 * lineage: an amplification adds copies of an allele (src_allele in the
 * reference), a deletion removes every copy of the allele in the region, a
 * whole-genome doubling doubles every copy; the copies of a lineage carry
 * its mutations. */
enum { PSX_CNA_AMP = 0, PSX_CNA_DEL = 1, PSX_CNA_WGD = 2 };
#define PSX_ALL_CHROMOSOMES 0xFFFFFFFFu
#define PSX_MAX_CNA_PLOIDY 8u

#define PSX_CNA_RANDOM_ALLELE 0xFFu

typedef struct psx_cna {
    uint32_t species;  /* on the edge into every node where this species first appears on a lineage */
    uint32_t chr;      /* ignored for PSX_CNA_WGD */
    uint32_t start;    /* 0-based first base */
    uint32_t len;      /* bases */
    uint8_t allele;    /* allele lineage (AMP, DEL); PSX_CNA_RANDOM_ALLELE: uniform per placement node */
    uint8_t kind;      /* PSX_CNA_AMP / PSX_CNA_DEL / PSX_CNA_WGD */
    uint8_t copies;    /* AMP: copies added */
    uint8_t pad;
} psx_cna;

/* A placed copy-number event. Events are kept in DFS order of their node
 * (an ancestor's events precede its descendants'), ties in driver order. */
typedef struct psx_cna_event {
    uint32_t node;     /* DFS index of the node whose incoming edge carries the event */
    uint32_t leaf_lo;  /* applies to leaves [leaf_lo, leaf_hi) */
    uint32_t leaf_hi;
    uint32_t chr;      /* PSX_ALL_CHROMOSOMES for a whole-genome doubling */
    uint32_t start;    /* [start, end) */
    uint32_t end;
    uint8_t allele;
    uint8_t kind;
    uint8_t copies;
    uint8_t pad;
} psx_cna_event;

/* A region of the passenger CNA catalogue (load_passenger_CNAs,
 * src/mutation_engine.cpp:441-487: a row with major copy number > 1 or NA
 * gives an amplification, one with minor copy number < 1 or NA a deletion). */
typedef struct psx_cna_region {
    uint32_t chr;
    uint32_t start;    /* 0-based first base */
    uint32_t len;
    uint8_t kind;      /* PSX_CNA_AMP or PSX_CNA_DEL */
    uint8_t pad[3];
} psx_cna_region;

/* COSMIC indel (ID83) classes, in COSMIC order: 1 bp deletions of C, then of
 * T, at homopolymers of 1..6+ bases (12); 1 bp insertions of C, then of T,
 * next to 0..5+ copies (12); deletions of 2, 3, 4, 5+ bp units at 1..6+
 * copies (24); insertions of 2, 3, 4, 5+ bp units next to 0..5+ copies
 * (24); deletions of 2, 3, 4, 5+ bp with microhomology of 1, 1-2, 1-3, 1-5
 * bases (11). */
#define PSX_ID_CLASSES 83u

typedef struct psx_engine_params {
    uint32_t n_species;
    const double* snv_rate;     /* [n_species] SNVs per base per division */
    const double* indel_rate;   /* [n_species] indels per base per division */
    uint32_t n_sbs;             /* SBS signatures, 96 classes each, COSMIC order */
    const double* sbs;          /* [n_sbs][96] */
    uint32_t n_epochs;          /* timed exposures (MutationEngine$add_exposure) */
    const double* epoch_start;  /* [n_epochs], ascending */
    const double* exposure;     /* [n_epochs][n_sbs] */
    uint32_t n_preneoplastic_snv;
    uint32_t n_preneoplastic_indel;
    uint32_t preneoplastic_sbs; /* signature index of the preneoplastic SNVs */
    uint32_t indel_max_len;     /* without ID signatures: indel lengths are geometric on [1, indel_max_len] */
    double indel_len_p;
    int32_t infinite_sites;     /* MutationEngine$infinite_sites_model */
    uint32_t ploidy;            /* copies of every chromosome (2) */
    uint32_t n_drivers;
    const psx_driver* drivers;
    uint32_t n_germline;
    const psx_germline* germline;
    uint32_t n_cna;             /* copy-number drivers (ploidy <= PSX_MAX_CNA_PLOIDY) */
    const psx_cna* cna;
    /* passenger CNAs: per edge Poisson(cna_rate(species) x ploidy x genome
     * length) events, each a uniform catalogue region on a uniform allele
     * (an amplification adds one copy of it); cna_rate NULL = none */
    const double* cna_rate;     /* [n_species] passenger CNAs per base per division */
    uint32_t n_cna_catalogue;
    const psx_cna_region* cna_catalogue;
    /* indel signatures (n_id > 0): an indel's ID class is drawn from the ID
     * exposure of its node's epoch, its site uniformly among the reference
     * sites of the class (classes without a site in the reference are left
     * out and the rest renormalised); n_id == 0 keeps geometric-length indels
     * at uniform positions with no cause */
    uint32_t n_id;              /* ID signatures, PSX_ID_CLASSES classes each */
    const double* id;           /* [n_id][PSX_ID_CLASSES] */
    const double* id_exposure;  /* [n_epochs][n_id] */
    uint32_t preneoplastic_id;  /* signature index of the preneoplastic indels */
    /* SBS sites: the genome's context index keeps every context_sampling-th
     * site (in genome order, from the first) of each of the 32
     * pyrimidine-centred trinucleotide contexts (MutationEngine's
     * context_sampling, src/mutations.cpp:818-819; 0 or 1 = every site); an
     * SNV's site is uniform over its class's indexed sites other than the
     * driver loci, and classes without an indexed site are left out of the
     * signatures and the rest renormalised. The reference must be shorter
     * than 2^32 bases. */
    uint32_t context_sampling;
} psx_engine_params;

/* One record per placed mutation; the table is sorted by SID
 * (chr, pos, kind, length, alternative sequence), ties by creation order. */
typedef struct psx_mutation {
    uint32_t chr;
    uint32_t pos;       /* 0-based */
    uint32_t leaf_lo;   /* carried by leaves [leaf_lo, leaf_hi) */
    uint32_t leaf_hi;
    uint32_t node;      /* DFS index of the edge node, PSX_NO_NODE for preneoplastic/germline */
    uint32_t seq;       /* packed bases: deleted reference bases (DEL), inserted bases (INS) */
    uint16_t cause;     /* SBS signature index, PSX_NO_CAUSE if none */
    uint8_t allele;
    uint8_t kind;       /* PSX_SNV / PSX_DEL / PSX_INS */
    uint8_t ref;        /* reference base at pos */
    uint8_t alt;        /* SNV alternative base */
    uint8_t len;        /* indel length */
    uint8_t nature;     /* PSX_PASSENGER / PSX_DRIVER / PSX_PRENEOPLASTIC / PSX_GERMINAL */
    uint8_t sid_head;   /* 1 on the first record of each SID */
    uint8_t pad[3];
} psx_mutation;

psx_status psx_place_mutations(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                               const psx_engine_params* p, uint64_t seed, psx_mutations** out);
/* Placement of n_jobs forests over one genome in one call, their device
 * phases run together with one synchronisation per phase. No counterpart in
 * the reference (which places one forest per MutationEngine call,
 * src/mutation_engine.cpp:1085-1113): the batched form for workloads of many
 * independent forests. Table k is identical to psx_place_mutations(ctx, g,
 * f[k], p[k], seeds[k]); the tables share one device block, freed with the
 * last of them. On an error (its message names the job) no table is made. */
psx_status psx_place_mutations_batch(psx_ctx* ctx, const psx_genome* g, uint32_t n_jobs,
                                     const psx_forest* const* f, const psx_engine_params* const* p,
                                     const uint64_t* seeds, psx_mutations** out);
/* Replay: a caller-provided (e.g. exported) table; it is sorted by SID here. */
psx_status psx_mutations_upload(psx_ctx* ctx, const psx_forest* f, const psx_mutation* recs,
                                uint64_t n, uint32_t ploidy, psx_mutations** out);
psx_status psx_mutations_count(const psx_mutations* m, uint64_t* n);
/* Copy-number events of the table (placed, or set for a replay): *n gets
 * their number; up to cap are copied to ev (ev may be NULL). */
psx_status psx_mutations_cna(const psx_mutations* m, psx_cna_event* ev, uint32_t cap, uint32_t* n);
psx_status psx_mutations_set_cna(psx_ctx* ctx, psx_mutations* m, const psx_cna_event* ev, uint32_t n);
psx_status psx_mutations_download(psx_ctx* ctx, const psx_mutations* m, psx_mutation* recs,
                                  uint64_t n);
void psx_mutations_free(psx_mutations* m);
/* Checkpoint and resume of a table: psx_mutations_save writes the records (SID
 * order), the copy-number events, the chromosome lengths, the ploidy and the
 * leaf count to a file in this library's binary format (header, version,
 * checksum); psx_mutations_load reads one back into device memory, ready for
 * sequencing. Replaces PhylogeneticForest::save / load_phylogenetic_forest
 * (src/phylogenetic_forest.cpp:940-992, src/mutations.cpp:1453-1478) for the
 * mutation table, whose RACES archive format is not this library's; a missing
 * or non-regular file, a wrong format (magic, size, checksum, a record out of
 * range) and another format version are refused with the reference loader's
 * messages (src/phylogenetic_forest.cpp:958-986, src/utility.cpp:141-161). */
#define PSX_TABLE_FORMAT_VERSION 1u
psx_status psx_mutations_save(psx_ctx* ctx, const psx_mutations* m, const char* path);
psx_status psx_mutations_load(psx_ctx* ctx, const char* path, psx_mutations** out);
/* Asks the call running on ctx (from another thread) to stop at its next
 * work-chunk boundary with PSX_ECANCELLED ("cancelled"): psx_write_sam checks
 * between chromosomes and between chunks of reads, psx_autotune_seq between
 * trials. The call that observes the request consumes it; a request made
 * while none of those calls runs stops the next one at its first boundary.
 * No counterpart in the reference, whose only interruption point is a user
 * interrupt checked every 10,000 events of the spatial simulation
 * (src/simulation.cpp:28-50), outside this path. */
psx_status psx_ctx_cancel(psx_ctx* ctx);
/* Milliseconds of the placement kernels of the last psx_place_mutations. */
psx_status psx_ctx_last_place_ms(const psx_ctx* ctx, double* ms);

/* --------------------------------------------------------------- sequencing */

typedef struct psx_seq_params {
    double coverage;
    uint32_t read_size;
    uint32_t insert_size_mean;    /* 0 = single-end reads */
    uint32_t insert_size_stddev;
    double error_rate;            /* 0 = error-less Illumina sequencer */
    double purity;                /* fraction of tumour cells in every tumour sample */
    int32_t with_normal_sample;
    int32_t preneoplastic_in_normal;
    uint32_t n_chromosomes;       /* 0 = every chromosome of the genome */
    const uint32_t* chromosomes;
    uint32_t n_samples;           /* forest samples to sequence, in order; with samples == NULL */
    const uint32_t* samples;      /* every forest sample in index order (n_samples is ignored) */
} psx_seq_params;

/* Validates the paired-end insert-size model Binomial(t, p) with
 * p = 1 - sd^2/mean and t = floor(mean / p). */
psx_status psx_insert_size_check(psx_ctx* ctx, uint32_t mean, uint32_t stddev);

/* Number of output samples: requested samples, plus one normal sample. */
uint32_t psx_seq_num_samples(const psx_forest* f, const psx_seq_params* p);
/* Reads the call simulates (closed form of the sampling model for mutation
 * tables without copy-number events; with them the count is random and the
 * sequencing calls report it). */
psx_status psx_seq_num_reads(const psx_genome* g, const psx_forest* f, const psx_seq_params* p,
                             uint64_t* n_reads);
/* Reads a sequencing call of this table and seed simulates, copy-number
 * events included (their kept templates are counted on the GPU with the
 * call's own draws): the capacity psx_simulate_seq_reads needs. */
psx_status psx_seq_num_reads_table(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                   const psx_mutations* m, const psx_seq_params* p, uint64_t seed,
                                   uint64_t* n_reads);

/* Copy numbers. With copy-number events on a chromosome, a template draws
 * its source cell and one of kmax copy slots (kmax = the chromosome's largest
 * copy number over leaves and positions, at least the ploidy) and is kept
 * when the slot is one of the source's copies at the template start: reads
 * then come from allele lineages in proportion to their copies and depth
 * scales with the total copy number (coverage is that of a genome with
 * `ploidy` copies; the normal genome has `ploidy` copies everywhere, so
 * normal admixture is weighted by DNA quantity). Both mates read the
 * template's lineage. The number of templates drawn is scaled by
 * kmax / ploidy; without events the model is the plain one, draw for draw. */

/* Count-only sequencing (write_SAM = FALSE). Output arrays are host memory
 * of n_out_samples x n_mutations entries, sample-major: alt = reads showing
 * the SID's alternative allele, depth = reads covering its position, present
 * = 1 if some source genome of the sample carries the mutation. */
psx_status psx_simulate_seq(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                            const psx_mutations* m, const psx_seq_params* p, uint64_t seed,
                            int32_t* alt, int32_t* depth, uint8_t* present, uint64_t* n_reads);
/* Same, writing device buffers (inputs and outputs stay in HBM). */
psx_status psx_simulate_seq_device(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                   const psx_mutations* m, const psx_seq_params* p,
                                   uint64_t seed, int32_t* d_alt, int32_t* d_depth,
                                   uint8_t* d_present, uint64_t* n_reads);

/* psx_simulate_seq_device with no device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The launch
 * descriptors and work counters live in a fixed device array of
 * PSX_NOALLOC_SCRATCH_BYTES per GPU (one copy per device, held until the
 * process ends), so a call whose descriptors exceed it fails with
 * PSX_EINVAL (descriptors take about 6 MB per output sample of a 3 Gb
 * genome, 2 bytes per 1,000 bases per output sample), as does a
 * call whose mutation table has not been indexed for the genome's chromosome
 * lengths yet (placement indexes a table for its own genome; an allocating
 * sequencing call indexes it for another). Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU, from any context, race. Results
 * equal psx_simulate_seq_device's. */
#define PSX_NOALLOC_SCRATCH_BYTES (32u << 20)
psx_status psx_simulate_seq_device_noalloc(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                           const psx_mutations* m, const psx_seq_params* p,
                                           uint64_t seed, int32_t* d_alt, int32_t* d_depth,
                                           uint8_t* d_present, uint64_t* n_reads);

/* Read materialisation (write_SAM = TRUE analogue): every read as an 8-byte
 * header (start position; source leaf | allele << 26 | mate << 30 | has-error
 * << 31) plus ceil(read_size/16) words of 2-bit bases with the source's SNVs
 * and the sequencing errors applied. d_reads holds psx_read_words(p) uint32
 * per read, reads ordered by (sample, chromosome, stratum, template, mate);
 * on a chromosome with copy-number events only the kept templates have reads
 * (psx_seq_num_reads_table counts them). */
uint32_t psx_read_words(const psx_seq_params* p);
psx_status psx_simulate_seq_reads(psx_ctx* ctx, const psx_genome* g, const psx_forest* f,
                                  const psx_mutations* m, const psx_seq_params* p,
                                  uint64_t seed, uint32_t* d_reads, uint64_t capacity_reads,
                                  uint64_t* n_reads);

/* SAM output of simulate_seq(write_SAM = TRUE) (src/seq_simulation.cpp:536-564;
 * vignettes/sequencing.Rmd:137-180, 272-309): one file per selected
 * chromosome, "<filename_prefix><chromosome name>.sam", holding the reads of
 * every output sample, one read group per sample. In update mode a
 * chromosome whose file exists is written to the first free
 * "<filename_prefix><chromosome name>_<n>.sam" (n = 1, 2, ...); otherwise an
 * existing file is an error. Reads are those of psx_simulate_seq_reads, made
 * one chromosome at a time on the device and written from pinned chunks
 * (no full copy of the reads in host memory). Template names are
 * "<template_name_prefix><k>", k numbering templates in (output sample,
 * chromosome) order. Qualities (Phred+33): constant at
 * min(40, round(-10 log10 e)) (40 when e = 0); with random_quality, each
 * base's score is drawn from Philox(seed, read, base) uniformly within 5 of
 * that value, clipped to [2, 41]. */
typedef struct psx_sam_params {
    const char* output_dir;            /* must exist */
    const char* filename_prefix;       /* "chr_" */
    const char* template_name_prefix;  /* "r" */
    int32_t update;                    /* UPDATE mode (update_SAM) */
    int32_t random_quality;            /* BasicIlluminaSequencer$random_quality_scores */
    const char* const* sample_names;   /* one per output sample (psx_seq_num_samples) */
    const char* const* chr_names;      /* one per genome chromosome */
} psx_sam_params;
psx_status psx_write_sam(psx_ctx* ctx, const psx_genome* g, const psx_forest* f, const psx_mutations* m,
                         const psx_seq_params* p, uint64_t seed, const psx_sam_params* sam,
                         uint64_t* n_reads);

/* ------------------------------------------------------------- batched mode */

typedef struct psx_seq_job {
    const psx_genome* genome;
    const psx_forest* forest;
    const psx_mutations* mutations;
    psx_seq_params params;
    uint64_t seed;
    int32_t* d_alt;       /* device, psx_seq_num_samples x n_mutations */
    int32_t* d_depth;
    uint8_t* d_present;
} psx_seq_job;

typedef struct psx_launch_config {
    uint32_t block_threads;    /* 64..1024, multiple of 64 */
    uint32_t lds_records;      /* mutation records staged in LDS per block */
    uint32_t strata_per_item;  /* strata processed by one work item */
    uint32_t blocks_per_cu;    /* persistent blocks per compute unit */
    int32_t use_graph;         /* capture the launch sequence in a hipGraph */
} psx_launch_config;

/* All jobs in persistent launches pulling (job, sample, chromosome, strata)
 * items from device queues: one launch for the jobs whose reads cover few
 * mutation positions (errors checked per covered position) and one for the
 * others (errors drawn per template); the same counts either way. cfg = NULL
 * uses the tuned default. */
psx_status psx_simulate_seq_batch(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                  const psx_launch_config* cfg, uint64_t* n_reads);
/* Read materialisation of many jobs in one persistent launch (the batched
 * form of psx_simulate_seq_reads; src/seq_simulation.cpp:536-549,564 writes
 * one SAM set per simulate_seq call): job k's reads are rows
 * [job_first[k], job_first[k + 1]) of d_reads, each job's rows exactly those
 * psx_simulate_seq_reads writes for it alone. Every job needs the same
 * psx_read_words; the jobs' d_alt, d_depth and d_present are not used.
 * job_first (host, n_jobs + 1 entries) may be NULL. */
psx_status psx_simulate_seq_reads_batch(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                                        uint32_t* d_reads, uint64_t capacity_reads, uint64_t* job_first,
                                        uint64_t* n_reads);
/* A batched launch prepared once and run many times: the descriptors are
 * built and uploaded at creation and, with cfg->use_graph (the default when
 * cfg is NULL), the launch is captured in a hipGraph that every run replays.
 * Each run rewrites the jobs' output arrays; the jobs' inputs and outputs
 * must outlive the plan. Runs on the context's stream; psx_ctx_last_kernel_ms
 * reports the replay (init and count kernels). */
typedef struct psx_seq_plan psx_seq_plan;
psx_status psx_seq_plan_create(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                               const psx_launch_config* cfg, psx_seq_plan** out);
psx_status psx_seq_plan_run(psx_seq_plan* plan, uint64_t* n_reads);
void psx_seq_plan_free(psx_seq_plan* plan);
/* Times candidate launch configurations on the jobs and returns the fastest
 * (results are identical for every configuration). */
psx_status psx_autotune_seq(psx_ctx* ctx, const psx_seq_job* jobs, uint32_t n_jobs,
                            psx_launch_config* best, double* best_ms);

/* ------------------------------------------------------ result table (host) */

/* get_result_dataframe semantics on the host arrays of psx_simulate_seq:
 * rows are the SIDs (sid_head records) covered by >= 1 alternative read in
 * some sample, or present in some sample when include_non_sequenced != 0.
 * row_rec[n_rows] receives record indices; occ/cov/vaf are
 * [n_out_samples][n_mutations]-capacity, filled [sample][row]. VAF is
 * occ/cov when the sample holds the SID (NaN if cov == 0), else 0. */
psx_status psx_seq_table(const psx_mutation* recs, uint64_t n_mutations,
                         uint32_t n_out_samples, const int32_t* alt, const int32_t* depth,
                         const uint8_t* present, int32_t include_non_sequenced,
                         uint64_t* n_rows, uint64_t* row_rec, int32_t* occ, int32_t* cov,
                         double* vaf);

/* The VAF column of get_result_dataframe (src/seq_simulation.cpp:123-130) in
 * fp32 on the device, for every record of the device arrays of
 * psx_simulate_seq_device ([n_out_samples][n_mutations]): at a SID head,
 * alt / depth (IEEE fp32 division) when the sample holds the SID (an
 * alternative read, or a record of the SID present in the sample; NaN when
 * depth is 0), else 0; at the other records of a SID, 0. d_vaf is device
 * memory of n_out_samples x n_mutations floats. */
psx_status psx_seq_vaf_f32(psx_ctx* ctx, const psx_mutations* m, uint32_t n_out_samples, const int32_t* d_alt,
                           const int32_t* d_depth, const uint8_t* d_present, float* d_vaf);

/* ------------------------------------------------------------------ serving */

/* A request loop over device-resident genomes, forests and mutation tables
 * (no counterpart in the reference, whose simulate_seq runs one call at a
 * time, src/seq_simulation.cpp:517-601). Requests are admitted in arrival
 * order while their output arrays fit in a fixed HBM budget, allocated once
 * at creation; up to max_inflight launches run at once on their own streams,
 * each taking every admitted request that waits (up to max_batch) into one
 * persistent launch, so requests that arrive while others run start as soon
 * as a launch slot and memory are free. Results are identical to
 * psx_simulate_seq on the same inputs, alone or in any mix. A genome, forest
 * and mutation table stay owned by the caller and must outlive the requests
 * that use them; a mutation table is used with one genome at a time. */
typedef struct psx_server psx_server;

typedef struct psx_server_config {
    uint64_t hbm_budget;       /* bytes for request outputs (alt, depth, present) */
    uint32_t max_inflight;     /* concurrent launches, 1..8 */
    uint32_t max_batch;        /* requests per launch, >= 1 */
    psx_launch_config launch;  /* kernel configuration of every launch (use_graph ignored) */
    /* a free lane starts a launch while another lane's launch runs only once
     * that launch has run overlap_after_us microseconds, or when max_batch
     * requests are queued (0: at once). Short launches then finish alone
     * instead of sharing the compute units with the next request's. */
    uint32_t overlap_after_us;
} psx_server_config;

typedef struct psx_request {
    const psx_genome* genome;
    const psx_forest* forest;
    const psx_mutations* mutations;
    psx_seq_params params;     /* the chromosome and sample lists are copied */
    uint64_t seed;
    int32_t* alt;              /* host, n_out_samples x n_mutations; NULL skips the copy */
    int32_t* depth;
    uint8_t* present;
} psx_request;

typedef struct psx_request_stats {
    psx_status status;         /* of the request: PSX_OK, or why it was not run */
    uint64_t n_reads;
    uint32_t batch_size;       /* requests in its launch */
    uint32_t lane;             /* launch slot that ran it */
    double queued_ms;          /* submission -> launch */
    double latency_ms;         /* submission -> results in host memory */
} psx_request_stats;

typedef struct psx_server_totals {
    uint64_t submitted, completed, failed, launches, reads;
    uint64_t budget_bytes, peak_bytes;  /* arena size and its high-water mark */
} psx_server_totals;

psx_status psx_server_create(psx_ctx* ctx, const psx_server_config* cfg, psx_server** out);
/* Validates and queues a request (PSX_EINVAL for invalid parameters,
 * PSX_ENOMEM_DEVICE when its outputs alone exceed the budget). */
psx_status psx_server_submit(psx_server* s, const psx_request* r, uint64_t* ticket);
/* Blocks until the request has finished; returns its status and forgets the
 * ticket. timeout_ms < 0 waits without limit; PSX_ECANCELLED on time-out
 * (the request stays queued). */
psx_status psx_server_wait(psx_server* s, uint64_t ticket, double timeout_ms, psx_request_stats* st);
psx_status psx_server_totals_get(psx_server* s, psx_server_totals* t);
const char* psx_server_last_error(psx_server* s);
/* Finishes the running launches, cancels queued requests, frees the budget. */
void psx_server_destroy(psx_server* s);

/* -------------------------------------------------------------- measurement */

/* Vector-ALU issue ceiling of the context's GPU (no counterpart in the
 * reference; the count-only sequencing kernel is bound by vector instruction
 * issue, not by HBM, and is reported against this). Runs one launch of
 * `iters` loop iterations per lane over every compute unit (256-thread
 * blocks, 8 per CU) and returns its HIP-event time; lane_iters (nullable)
 * gets lanes x iters. PSX_PROBE_PHILOX iterations draw one Philox4x32-10
 * block each, PSX_PROBE_INT iterations run eight independent integer
 * add/shift/xor chains. */
enum { PSX_PROBE_PHILOX = 0, PSX_PROBE_INT = 1 };
psx_status psx_probe_valu(psx_ctx* ctx, int mode, uint32_t iters, double* ms, uint64_t* lane_iters);

/* HBM write ceiling of the context's GPU (no counterpart in the reference;
 * the read-materialising kernel is bound by its output stores and is
 * reported against this). One launch of stores only into `out` (device
 * memory of `bytes`, 16-byte aligned, contents overwritten), timed with HIP
 * events; *written gets the bytes stored. PSX_PROBE_STORE_FILL: grid-stride,
 * consecutive lanes on consecutive 16-byte words. PSX_PROBE_STORE_CHUNKS: the
 * read kernel's write shape, items of 170 chunks of 3 KiB taken by 512-thread
 * blocks from a counter, 2 chunks per wave per round, chunks on 1024-byte
 * boundaries; _SHIFTED: the same 16 bytes off them. */
enum { PSX_PROBE_STORE_FILL = 0, PSX_PROBE_STORE_CHUNKS = 1, PSX_PROBE_STORE_CHUNKS_SHIFTED = 2 };
psx_status psx_probe_store(psx_ctx* ctx, void* out, uint64_t bytes, int mode, double* ms, uint64_t* written);

/* ------------------------------------------- synthetic forest (host input) */

/* Birth-death genealogy standing in for the CPU spatial simulation, which
 * stays outside the accelerated path. Species 0 founds the tumour; species
 * s > 0 is introduced by converting one uniformly chosen species-0 cell when
 * the population reaches intro_at[s] cells. Growth stops at n_cells; then
 * n_samples disjoint samples of sample_size cells are drawn uniformly. */
typedef struct psx_synth_params {
    uint32_t n_species;
    const double* growth;     /* [n_species] division rate */
    const double* death;      /* [n_species] death rate */
    const uint64_t* intro_at; /* [n_species], ignored for species 0 */
    uint64_t n_cells;
    uint32_t n_samples;
    uint64_t sample_size;
} psx_synth_params;

typedef struct psx_synth_forest psx_synth_forest;
psx_status psx_synth_forest_run(const psx_synth_params* p, uint64_t seed, psx_synth_forest** out);
uint32_t psx_synth_forest_nodes(const psx_synth_forest* s);
/* Arrays of psx_synth_forest_nodes() entries, in creation order. */
psx_status psx_synth_forest_get(const psx_synth_forest* s, int64_t* cell_id, int64_t* parent,
                                uint32_t* species, int32_t* sample, double* birth_time);
void psx_synth_forest_free(psx_synth_forest* s);

#ifdef __cplusplus
}
#endif
#endif /* PSX_H */
