/*
 * gsnp_abi.h -- C ABI of the MI355X SOAPsnp calling library (libgsnp.so).
 *
 * This is the boundary the reference's window driver (cpp/call_genotype.cc)
 * crosses into the GPU.  Each entry point names the reference interface it
 * replaces (paths relative to the reference root).  The reference passes
 * implicit process-global device state; here an explicit context owns all
 * device memory.  Every function returns GSNP_OK or an error code, with a
 * message from gsnp_last_error().  A context is used by one host thread at a
 * time (the reference is single-threaded and not reentrant either).
 *
 * Host pointers are plain caller-owned arrays; device memory is owned by the
 * context.  No torch or HIP types appear in the signatures.
 */
#ifndef GSNP_ABI_H
#define GSNP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes (the reference aborts instead; cpp/gsnp.h has no status codes) */
#define GSNP_OK 0
#define GSNP_E_IO 1
#define GSNP_E_FORMAT 2
#define GSNP_E_UNSORTED 3   /* "Errors in sorting", cpp/call_genotype.cc:516-519 */
#define GSNP_E_CHROM 4      /* "Chromosome NOT found", cpp/matrix.cc:71-78 */
#define GSNP_E_QUAL 5       /* unique-hit base quality outside [1,63] (record field overflow) */
#define GSNP_E_RANGE 6      /* read length, hit count, adjusted quality out of range */
#define GSNP_E_NOMEM 7
#define GSNP_E_ARG 8
#define GSNP_E_DEVICE 9     /* HIP runtime error */

#define GSNP_WIN 262144     /* sites per window, cpp/common.h:16 */

/* Calling parameters; mirrors Parameter, cpp/common.h:123-154 (defaults :142-153). */
typedef struct gsnp_params {
    int q_min;               /* char standing for quality 0 (-z), default 64 */
    int q_max;               /* q_min + 40 (-Q) */
    int read_length;         /* -L, default 45 */
    int is_monoploid;        /* -m */
    int refine_mode;         /* -2 */
    int rank_sum_mode;       /* -u */
    int transition_dominant; /* -t */
    double althom_novel_r;   /* -r */
    double het_novel_r;      /* -e */
    double althom_val_r;
    double het_val_r;
    double althom_unval_r;
    double het_unval_r;
    double global_dependency; /* log10 scale (-g value is log10'd, cpp/main.cc:151-155) */
    double pcr_dependency;    /* log10 scale (-p, cpp/main.cc:156-160) */
} gsnp_params;

/* Known-SNP entry; mirrors Snp_info, cpp/soap_snp.h:714-761. freq is A C T G. */
typedef struct gsnp_snp {
    uint8_t hapmap, validated, indel, pad;
    double freq[4];
} gsnp_snp;

/* A batch of alignments for one chromosome, structure of arrays; mirrors
 * st_update, cpp/gsnp.h:134-310.  Reads must be position sorted and already
 * filtered to pos >= 0 and pos + len < chromosome length
 * (cpp/call_genotype.cc:495-514).  seq/qual hold `pitch` chars per read. */
typedef struct gsnp_read_batch {
    uint64_t n;
    uint32_t pitch;
    const int32_t* pos;      /* 0-based */
    const int32_t* len_hit;  /* (len << 16) | hit, cpp/gsnp.h:199-202 */
    const char* strand;      /* '+' forward, anything else reverse */
    const char* seq;
    const char* qual;
} gsnp_read_batch;

/* Per-site results of one called window (host copies). */
typedef struct gsnp_window_out {
    uint8_t* base1;     /* may be NULL for any field not wanted */
    uint8_t* base2;
    uint8_t* type1;
    int32_t* q_cns;
    double* rank_sum;
    double* lk10;       /* 10 genotype log-likelihoods per site */
    int32_t* depth;
    int32_t* dep_uni;
} gsnp_window_out;

typedef struct gsnp_ctx gsnp_ctx;

void gsnp_params_default(gsnp_params* p);
const char* gsnp_last_error(void);

/* Select the HIP device.  Replaces GPUInit, cpp/gsnp.h:26 (kernel/gsnp_kernel.cu:20-22). */
int gsnp_device_init(int device);

/* Build the score table, adjusted-quality table, priors and rank-sum table
 * and allocate window buffers.  p_matrix is 256*256*16 doubles indexed
 * q<<12|cycle<<4|ref<<2|obs.  Replaces GSNPKernelInit, cpp/gsnp.h:127-129
 * (kernel/gsnp_kernel.cu:24-33). */
int gsnp_create(const gsnp_params* p, const double* p_matrix, gsnp_ctx** out);

/* Replace the score table with one built from another calibration matrix
 * (same layout as gsnp_create's).  Replaces the table half of
 * GSNPKernelInit, kernel/gsnp_kernel.cu:24-33, for a context whose matrix is
 * counted after its reads are on the device. */
int gsnp_set_matrix(gsnp_ctx* ctx, const double* p_matrix);

/* Calibration counts of the reads loaded into the context, added into
 * counts[64 << 12], bin (q - q_min) << 12 | cycle << 4 | ref << 2 | obs:
 * unique-hit bases observed at or above q_min on reference bases that are
 * neither N nor a known SNP.  Replaces the counting loop of
 * Prob_matrix::matrix_gen, cpp/matrix.cc:58-121. */
int gsnp_calib_count(gsnp_ctx* ctx, uint64_t* counts);

/* Replaces GSNPKernelFinalize, kernel/gsnp_kernel.cu:36-42. */
void gsnp_destroy(gsnp_ctx* ctx);

/* Queue the context's device work on the caller's HIP stream (a
 * hipStream_t, not the null stream), so that its calls are ordered against
 * the caller's own work and two contexts can run side by side; NULL goes back
 * to the stream the context created.  Work already queued is finished first.
 * Entry points that hand results to host memory still synchronize that stream
 * before they return.  The reference runs on CUDA's default stream and has no
 * counterpart (cpp/gsnp.h:375-516 takes no stream). */
int gsnp_set_stream(gsnp_ctx* ctx, void* stream);

/* Reference sequence of the chromosome to call, as 4-bit codes per base
 * (cpp/soap_snp.h:763-789: bit3 known SNP, bit2 N, bits1:0 A0 C1 T2 G3), and
 * its known-SNP entries (sorted positions).  Replaces the per-window host
 * gathers of get_bin_base and find_snp, cpp/call_genotype.cc:222-262. */
int gsnp_set_chromosome(gsnp_ctx* ctx, const char* name, const uint8_t* codes, uint32_t len,
                        const uint32_t* snp_pos, const gsnp_snp* snps, uint32_t n_snp);

/* Validate, pack and upload the chromosome's reads into device memory
 * (2-bit base + 6-bit quality per byte).  Replaces the per-window
 * ~218 MB st_update copy of cuda_updateSparse, kernel/update_kernel.cu:346-348. */
int gsnp_load_reads(gsnp_ctx* ctx, const gsnp_read_batch* batch);

/* Reset window state to start at `start`.  Replaces cuda_init, cpp/gsnp.h:492-498. */
int gsnp_window_init(gsnp_ctx* ctx, uint32_t start);

/* Add reads [r0, r1) of the loaded batch (all starting inside the current
 * window) to the site matrix; returns the largest record count per site.
 * Replaces gold_pos_match + cuda_updateSparse, cpp/gsnp.h:385, :403-417. */
int gsnp_window_update(gsnp_ctx* ctx, uint64_t r0, uint64_t r1, uint32_t* max_records);

/* Segmented sort of each site's records and genotype likelihoods for sites
 * [0, call_length).  Replaces cuda_likelihoodSparse, cpp/gsnp.h:398-401. */
int gsnp_window_likelihood(gsnp_ctx* ctx, uint32_t call_length);

/* Posterior, best/second genotype, consensus quality and (-u) rank-sum test.
 * Replaces cuda_posterior, cpp/gsnp.h:470-489. */
int gsnp_window_posterior(gsnp_ctx* ctx, uint32_t call_length);

/* Format the window's .cns rows in device memory; *nbytes receives their
 * size.  Replaces the text branch of Call_win::call_cns2,
 * cpp/call_genotype.cc:353-411. */
int gsnp_window_text(gsnp_ctx* ctx, uint32_t call_length, uint64_t* nbytes);

/* Copy the last formatted rows (gsnp_window_text / gsnp_batched_call) to host. */
int gsnp_fetch_text(gsnp_ctx* ctx, char* dst, uint64_t cap, uint64_t* nbytes);

/* The -C chunks of the window just called (gsnp_window_posterior), encoded
 * on the device: sizes[f] receives the bytes of file f's chunk (.3_4_6,
 * .5_7_8_9_14_16, .10, .11_12_13, .15_17) and, with dst non-NULL, dst[f]
 * (cap[f] bytes) the chunk itself; a buffer too small fails with GSNP_E_ARG
 * after sizes is filled.  call_length 0 gives the chunks of an empty window.
 * The window's .1_2 row is (chromosome, start + 1, call_length).  Replaces
 * the -C branch of Call_win::call_cns2, cpp/call_genotype.cc:312-352. */
int gsnp_window_columns(gsnp_ctx* ctx, uint32_t call_length, uint8_t* const* dst, const uint64_t* cap,
                        uint64_t* sizes);

/* Copy per-site results of the last called window to host. */
int gsnp_window_results(gsnp_ctx* ctx, uint32_t call_length, const gsnp_window_out* out);

/* Carry the read-length tail to the front of the next window.  Replaces
 * cuda_recycle / cuda_quickRecycle, cpp/gsnp.h:499-516. */
int gsnp_window_recycle(gsnp_ctx* ctx);

/* Batched form: call sites [start, start+n_sites) of the chromosome straight
 * from the device read store in one fused launch per depth bucket, without
 * the per-window recycle; emit_text formats the rows.  Same outputs as the
 * per-window path (no reference counterpart; it replaces the window loop of
 * Call_win::new_soap2cns, cpp/call_genotype.cc:474-644). */
int gsnp_batched_call(gsnp_ctx* ctx, uint32_t start, uint32_t n_sites, int emit_text, uint64_t* nbytes);

/* Autotuner of the batched form: builds the span's site matrix, then times
 * the fused launch of every depth class on it for each smallest network width
 * (8, 16, 32: the lowest bucket edge), waves per block (1, 2, 4, 8) and grid cap
 * (512..8192 blocks), and keeps the fastest shape in the context for later
 * gsnp_batched_call / gsnp_job_call.  reps timings per candidate, best kept.
 * config (13 entries, may be NULL) receives min class, waves per block of the
 * six classes, grid caps of the six classes.  No reference counterpart. */
int gsnp_autotune(gsnp_ctx* ctx, uint32_t start, uint32_t n_sites, int reps, uint32_t* config);

/* Whole pipeline on files: FASTA + sorted SOAP alignments (+ known SNPs) to
 * .cns text or, with compress, the -C column files <out>.1_2 ... <out>.15_17.
 * matrix_in/matrix_out are the -I/-M calibration files (may be NULL).
 * batched selects gsnp_batched_call instead of the per-window path.
 * Replaces main + Call_win::new_soap2cns, cpp/main.cc:91-429. */
int gsnp_run(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
             const gsnp_params* p, const char* out, int compress, int batched);

/* -C files back to .cns text.  Replaces decompressAll, cpp/decompression.cpp:130-255. */
int gsnp_decompress(const char* prefix, const char* out);
/* The same on the current HIP device: the host reads and checks each
 * window's chunks; the device rebuilds the 17 columns (2-bit words,
 * dictionary indices, run lengths, sparse lists) and writes the rows, up to
 * 16 windows per upload. */
int gsnp_decompress_device(const char* prefix, const char* out);

/* -C files read in memory one window at a time, as the 17 .cns columns.
 * Replaces CNSDataWindow and the read_decompress loop of decompressAll,
 * cpp/decompression.cpp:7-255 (the in-memory extraction the reference's
 * README describes).  The arrays belong to the reader and stay valid until
 * its next gsnp_cns_next or gsnp_cns_close. */
typedef struct gsnp_cns_reader gsnp_cns_reader;
typedef struct {
    const char* chr;       /* column 1 */
    uint64_t start;        /* column 2 of row 0 (1-based) */
    uint64_t n;            /* rows */
    const char* ref;       /* column 3 */
    const char* cns;       /* column 4 */
    const int32_t* q_cns;  /* column 5 */
    const char* base1;     /* column 6 */
    const int32_t* avg_q1, *count_uni1, *count_all1;  /* columns 7, 8, 9 */
    const char* base2;                                 /* column 10 */
    const int32_t* avg_q2, *count_uni2, *count_all2;  /* columns 11, 12, 13 */
    const int32_t* depth;                              /* column 14 */
    const double* rank_sum, *copy_num;                 /* columns 15, 16 */
    const int32_t* dbsnp;                              /* column 17 */
} gsnp_cns_window;
int gsnp_cns_open(const char* prefix, gsnp_cns_reader** out);
/* *has_window = 0 once every window has been read */
int gsnp_cns_next(gsnp_cns_reader* reader, gsnp_cns_window* w, int* has_window);
void gsnp_cns_close(gsnp_cns_reader* reader);

/* Counting-trained calibration matrix from a SOAP file, FASTA and optional
 * known-SNP file (whose sites are excluded from counting); the counts are
 * taken on the device (gsnp_calib_count).  Replaces Prob_matrix::matrix_gen,
 * cpp/matrix.cc:42-213. */
int gsnp_matrix_gen(const char* fasta, const char* soap, const char* dbsnp, const gsnp_params* p, double* p_matrix);

/* Dictionary encoding of a 32-bit integer column on the device: sorted
 * unique dictionary plus MSB-first bit packing.  Replaces cuda_dictCompression
 * (int overload), cpp/dict_compression.h:26-31.  words needs n entries, dict n. */
int gsnp_dict_compress_i32(gsnp_ctx* ctx, const int32_t* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                           int32_t* dict, uint64_t* n_dict);
/* The char, uint32 and double overloads of cuda_dictCompression,
 * cpp/dict_compression.h:26-37 (kernel/compress_kernel.cu:270-334): the
 * same layout, the dictionary in the value type's order (char signed).  For
 * double, +0.0 and -0.0 are one value (the dictionary holds +0.0 unless the
 * input has only -0.0) and NaN is rejected (GSNP_E_ARG).  n <= 2^30. */
int gsnp_dict_compress_char(gsnp_ctx* ctx, const char* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                            char* dict, uint64_t* n_dict);
int gsnp_dict_compress_u32(gsnp_ctx* ctx, const uint32_t* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                           uint32_t* dict, uint64_t* n_dict);
int gsnp_dict_compress_f64(gsnp_ctx* ctx, const double* in, uint64_t n, uint32_t* words, uint64_t* n_words,
                           double* dict, uint64_t* n_dict);

/* Segmented sort + likelihood on caller-supplied records: n_sites rows of
 * `pitch` u32 records, counts[i] valid in row i; writes 10 log-likelihoods
 * per site.  Exposes the sort and likelihood kernels on their own
 * (cpp/gsnp.h:387-401). */
int gsnp_sort_likelihood(gsnp_ctx* ctx, uint32_t* records, uint32_t pitch, const uint32_t* counts,
                         uint32_t n_sites, double* lk10);

/* No counterpart in the reference.  gsnp_sort_likelihood on caller-owned
 * device memory, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): it calls no
 * hipMalloc, hipFree or other allocating routine.  Its working memory (depth
 * class lists, class counters, an error word) is a fixed device array of
 * 6 MiB + 256 B per GPU, held until the process ends; its device address is
 * looked up by gsnp_device_init (or by the first call on a device).
 * records: device, n_sites rows of 256 u32; the first counts[i] records of
 * row i are sorted in place.  counts: device, n_sites int32 in [0, 256].
 * lk10: device, 10 doubles per site.  n_sites <= GSNP_WIN (262144), else
 * GSNP_E_ARG.  Runs on the context's stream and returns once it has drained.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (two host threads, two contexts) race and must be serialised by the caller. */
int gsnp_sort_likelihood_noalloc(gsnp_ctx* ctx, uint32_t* records, const int32_t* counts, uint32_t n_sites,
                                 double* lk10);

/* Per-stage device time from HIP events recorded on the context's stream:
 * returns (and resets) the accumulated milliseconds and launch counts of the
 * stages site build, sort, likelihood, posterior, rank-sum, format, recycle,
 * fused (the batched form's sort + likelihood + posterior launches), adjoint,
 * -C columns, calibration counts, depth-class bucketing of a span (12
 * entries), then turns
 * recording on or off.  Replaces the cutil
 * timers of the reference (kernel/cuda_header.h:37-50). */
#define GSNP_N_STAGES 12
int gsnp_profile(gsnp_ctx* ctx, int enable, double* stage_ms, uint64_t* stage_count, int n_stages);

/* Observation records the likelihood stage has consumed since the last
 * reset (throughput accounting). */
int gsnp_stats(gsnp_ctx* ctx, uint64_t* n_records, int reset);

/* Counters since the last reset, the first n of: [0] observation records
 * called, [1] sites finished on the host with the host libm (a rank-sum
 * normal approximation whose exp or sqrt lies too close to a rounding
 * boundary, or a consensus quality within the guard band of an integer),
 * [2] those of [1] that took the normal approximation (n1 + n2 >= 64),
 * [3] normal approximations the device finished. */
int gsnp_counters(gsnp_ctx* ctx, uint64_t* out, int n, int reset);

/* Session over input files: parses FASTA, SOAP and known SNPs, trains (or
 * reads) the calibration matrix and creates the context, so that calling
 * passes can be repeated on device-resident inputs.  Replaces the pass-1
 * set-up of main, cpp/main.cc:363-403. */
typedef struct gsnp_job gsnp_job;
int gsnp_job_open(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in,
                  const gsnp_params* p, gsnp_job** out);
/* number of chromosomes with alignments, in call order */
int gsnp_job_count(gsnp_job* job);
int gsnp_job_info(gsnp_job* job, int index, char* name, uint64_t cap, uint32_t* length, uint64_t* n_reads);
/* upload a chromosome's reference and reads to the device */
int gsnp_job_select(gsnp_job* job, int index);
/* one full calling pass over the selected chromosome (batched = windows per
 * fused span, 0 for the per-window path).  emit_text 0: no rows; 1: rows
 * formatted in device memory; 2: rows also copied to pinned host memory (per
 * window with one synchronous copy; per span with the copy of each span
 * running beside the next span's kernels).  *nbytes sums the rows.  Replaces
 * the window loop of new_soap2cns, cpp/call_genotype.cc:474-644. */
int gsnp_job_call(gsnp_job* job, int batched, int emit_text, uint64_t* nbytes);
/* the same over every chromosome of the job in call order (a whole-genome
 * pass; chromosomes stay resident on the device once uploaded, so a pass
 * moves no inputs); batched < 0 makes each chromosome one span.  emit_text 3
 * encodes the -C column chunks on the device and copies them to host memory
 * instead of rows (*nbytes then sums the chunk bytes).  *n_sites sums the
 * called sites. */
int gsnp_job_pass(gsnp_job* job, int batched, int emit_text, uint64_t* nbytes, uint64_t* n_sites);
/* the same pass with the -C column chunks of every window encoded on the
 * device and copied to host memory, but not written to files (the encoder
 * of gsnp_run with compress = 1); *nbytes sums the chunk bytes.  Replaces the
 * extraction and compression of Call_win::call_cns2 with -C,
 * cpp/compress.h:33-1312. */
int gsnp_job_columns(gsnp_job* job, int batched, uint64_t* nbytes);
/* The same pass handing every window's five chunks to fn, in window order:
 * chunk[f] / bytes[f] for the files .3_4_6, .5_7_8_9_14_16, .10, .11_12_13,
 * .15_17, and the window's .1_2 row (chr, 1-based start, n_sites).  The
 * chunks are valid during the call only; a nonzero return from fn ends the
 * pass with GSNP_E_IO.  *nbytes sums the chunk bytes.  Replaces the -C
 * branch of Call_win::call_cns2 (cpp/call_genotype.cc:312-352) for a whole
 * chromosome. */
typedef int (*gsnp_chunk_fn)(void* user, const char* chr, uint64_t start, uint32_t n_sites,
                             const uint8_t* const* chunk, const uint64_t* bytes);
int gsnp_job_columns_to(gsnp_job* job, int batched, gsnp_chunk_fn fn, void* user, uint64_t* nbytes);
/* Every chromosome of the job called and written to `out` (.cns text) or,
 * with compress, to the -C files <out>.1_2 ... <out>.15_17; batched as in
 * gsnp_run.  The calling and output half of main (cpp/main.cc:391-406). */
int gsnp_job_write(gsnp_job* job, const char* out, int compress, int batched);
/* -T target regions (Genome::read_region and Chr_info::set_region,
 * cpp/chromosome.cc:66-170; the window gating of cpp/call_genotype.cc:535,
 * :593): lines "chr start end" (1-based, inclusive) select the windows that
 * [start - read_length, end - 1] touches; later calls and writes of the job
 * call and output only those windows (a chromosome without a line: none),
 * each built from every read that covers it.  A null or empty path clears
 * the regions.  Lines of unknown chromosomes are skipped. */
int gsnp_job_set_regions(gsnp_job* job, const char* path);
/* The job's calibration matrix as the -M text file (cpp/matrix.cc:230-241). */
int gsnp_job_write_matrix(gsnp_job* job, const char* path);
/* A job over a seeded synthetic data set generated in memory (the data
 * gsnp_synth_write writes as files, for configurations without indel edits:
 * 1-5), calibrated on the device as gsnp_job_open does (or from matrix_in).
 * p's read length is replaced by the configuration's (*read_len).  No
 * reference counterpart: it feeds benchmarks without text files. */
int gsnp_job_synth(int config, unsigned long long seed, double scale, const gsnp_params* p, const char* matrix_in,
                   gsnp_job** out, int* read_len);
gsnp_ctx* gsnp_job_context(gsnp_job* job);
void gsnp_job_close(gsnp_job* job);

/* Adjoint of the calling likelihood (no reference counterpart; the reference
 * computes no gradient).  For the sites [start, start+n_sites) of the selected
 * chromosome, with lk_s(g) the site's genotype log10-likelihoods and prior_s
 * the calling priors (cpp/prior.cc), the objective is
 *   l = sum over called sites of ln sum_g prior_s(g) 10^lk_s(g).
 * gsnp_adjoint_prepare builds the span's records, their dependency-adjusted
 * qualities and the key order used for the deterministic reduction (once per
 * span).  gsnp_adjoint_eval takes a p_matrix (256*256*16 doubles, layout
 * q<<12|cycle<<4|ref<<2|obs) and ln het_rate, and returns l, dl/d ln p for
 * q < 64 (64*256*16 doubles, same layout; may be NULL), dl/d ln het and
 * optionally the per-site terms (n_sites doubles).  The score rows are
 * evaluated on the device from p_matrix; sums are in a fixed order, so
 * repeated calls return identical bits. */
int gsnp_adjoint_prepare(gsnp_ctx* ctx, uint32_t start, uint32_t n_sites);
int gsnp_adjoint_eval(gsnp_ctx* ctx, const double* p_matrix, double log_het, double* loglik, double* grad_log_p,
                      double* grad_log_het, double* site_loglik);
int gsnp_adjoint_info(gsnp_ctx* ctx, uint32_t* n_sites, uint64_t* n_records, uint32_t* n_prior_rows);
/* The prepared span's inputs to the objective: keys[r] = q<<10|cycle<<2|o of
 * each record (site order), offsets[n_sites+1] of each site's records, and
 * prior_row[s] (0..7 reference-code rows, 8+i known SNP i, 0xFFFFFFFF for an
 * uncalled N site).  Any pointer may be NULL. */
int gsnp_adjoint_export(gsnp_ctx* ctx, uint32_t* keys, uint64_t* offsets, uint32_t* prior_row);
/* log10 prior rows at ln het = log_het (10 per row, genotype order AA AC AT
 * AG CC CT CG TT TG GG) and per row a 10-bit mask of the genotypes whose
 * prior is proportional to het. */
int gsnp_adjoint_priors(gsnp_ctx* ctx, double log_het, double* lp10, uint32_t* het_mask, uint32_t cap_rows);

/* %#g text of one double as written in .cns rows (host build of the device
 * formatter). Returns the length. */
int gsnp_format_g(double v, char* out);

/* Seeded synthetic inputs: <prefix>.fa, .soap (and .snp for config 2).
 * Returns the number of alignment lines. */
long long gsnp_synth_write(int config, unsigned long long seed, double scale, const char* prefix, int* read_len);
/* The same files for the listed contigs only (indices in generation order). */
long long gsnp_synth_write_contigs(int config, unsigned long long seed, double scale, const char* prefix,
                                   const int* contigs, int n_contigs, int* read_len);
/* Contig count of a configuration; the first `cap` lengths go to lens (may be NULL). */
int gsnp_synth_contigs(int config, unsigned long long seed, double scale, uint64_t* lens, int cap);

#ifdef __cplusplus
}
#endif
#endif
