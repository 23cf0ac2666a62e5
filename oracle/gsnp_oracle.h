/*
 * gsnp_oracle.h -- CPU restatement of GSNP's SOAPsnp calling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  Nothing under
 * gsnp_amd/ links, includes or calls it.
 *
 * Every function cites the reference file:line it restates (paths are
 * relative to the reference repository root).  The reference ships no
 * golden vectors; the pins for each function are listed in the header of
 * gsnp_oracle.c and exercised by tests/test_oracle_*.py.
 *
 * Plain C99, strict IEEE double (built with -ffp-contract=off, no
 * -ffast-math), glibc libm for every transcendental.
 */
#ifndef GSNP_ORACLE_H
#define GSNP_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_WIN 262144          /* cpp/common.h:16 global_win_size = 32*8192 */
#define ORC_MAX_REC 256         /* cpp/common.h:15 MAX_NUM_NON_ZERO */
#define ORC_PMATRIX (256 * 256 * 16) /* cpp/common.h:81 PMATRIX_SIZE */
#define ORC_PRANK (64 * 64 * 2048)   /* cpp/matrix.cc:9 */
#define ORC_SCORE (256 * 256 * 4 * 10) /* kernel/likelihood_kernel.cu:9 */

/* error codes shared with the product's C ABI (same numbers) */
enum {
    ORC_OK = 0,
    ORC_E_IO = 1,
    ORC_E_FORMAT = 2,
    ORC_E_UNSORTED = 3,
    ORC_E_CHROM = 4,
    ORC_E_QUAL = 5,
    ORC_E_RANGE = 6,
    ORC_E_NOMEM = 7,
    ORC_E_ARG = 8
};

/* cpp/common.h:123-154 Parameter, defaults at :142-153 */
typedef struct {
    int q_min;               /* -z, char standing for quality 0, default '@' (64) */
    int q_max;               /* q_min + 40 (-Q), default 104 */
    int read_length;         /* -L, default 45 */
    int is_monoploid;        /* -m */
    int refine_mode;         /* -2 */
    int rank_sum_mode;       /* -u */
    int transition_dominant; /* -t */
    double althom_novel_r;   /* -r, default 0.0005 */
    double het_novel_r;      /* -e, default 0.0010 */
    double althom_val_r;     /* -a... see main.cc:194-213 */
    double het_val_r;
    double althom_unval_r;
    double het_unval_r;
    double global_dependency; /* log10 scale, default log10(0.9) */
    double pcr_dependency;    /* log10 scale, default log10(0.5) */
} orc_params;

void orc_params_default(orc_params* p);

/* ------------------------------------------------------------------ genome */
typedef struct {
    uint8_t hapmap, validated, indel;
    double freq[4];
} orc_snp;

typedef struct {
    char* name;
    uint32_t len;
    uint8_t* code;      /* one 4-bit code per base (bit3 dbSNP, bit2 N, bits1:0 A0 C1 T2 G3) */
    int nsnp;
    uint32_t* snp_pos;  /* sorted, unique (first insertion wins) */
    orc_snp* snp;
} orc_chr;

typedef struct {
    int n;
    orc_chr* chr;
} orc_genome;

int orc_genome_load(const char* fasta, const char* dbsnp, orc_genome** out);
void orc_genome_free(orc_genome* g);
int orc_genome_find(const orc_genome* g, const char* name);
const orc_snp* orc_find_snp(const orc_chr* c, uint32_t pos);

/* ------------------------------------------------------------------- reads */
typedef struct {
    int64_t n;
    int L;              /* pitch = read_length parameter */
    int32_t* pos;       /* 0-based (SOAP pos - 1) */
    int32_t* len;
    int32_t* hit;
    uint8_t* strand;    /* raw strand char */
    int32_t* chr;       /* index into genome, -1 if pos<0 and chromosome unknown */
    char* seq;          /* n*L raw base chars after indel edits ('N' padded) */
    char* qual;         /* n*L raw quality chars */
} orc_reads;

int orc_reads_load(const char* soap, const orc_params* p, const orc_genome* g, orc_reads** out);
void orc_reads_free(orc_reads* r);

/* --------------------------------------------------------------- tables */
int orc_matrix_gen(const orc_reads* r, const orc_genome* g, const orc_params* p, double* pm);
int orc_matrix_write(const char* path, const double* pm, const orc_params* p);
int orc_matrix_read(const char* path, double* pm);
void orc_prior_gen(const orc_params* p, double* p_prior);
void orc_rank_table_gen(double* p_rank);
void orc_score_table(const double* pm, double* score);

/* ------------------------------------------------ single-site primitives */
uint32_t orc_encode_word(uint32_t coord, uint32_t strand, uint32_t o_base, uint32_t q, uint32_t occ);
void orc_decode_word(uint32_t w, uint32_t* coord, uint32_t* strand, uint32_t* o_base, uint32_t* q, uint32_t* occ);
void orc_sort_records(uint32_t* rec, int n);
int orc_likelihood_sorted(const uint32_t* rec, int n, int read_length, double pcr, double glob,
                          const double* score, double* lk10);
int orc_likelihood_dense(const uint32_t* rec, int n, int read_length, int q_max_min, double pcr,
                         double glob, const double* score, double* lk10);
int orc_qadj(int q, int t, int g, double pcr, double glob);
double orc_normal_value(double z);
double orc_normal_test(int n1, int n2, double T1, double T2);
double orc_table_test(const double* p_rank, int n1, int n2, double T1, double T2);
double orc_rank_test(const uint32_t* rec, int n, const int* count_uni, int q_max_min, int best_type,
                     const double* p_rank);
int orc_snp_prior(double* row16, int ref, const orc_snp* s, const orc_params* p);
int orc_trunc_int(double x);
int orc_format_g(double v, char* out);

/* per-site result, filled by orc_call_window */
typedef struct {
    int32_t depth, repeat, dep_uni;
    int32_t cnt_uni[4], q_sum[4], cnt_all[4];
    uint8_t ori, nflag, base1, base2, type1, type2;
    int32_t qual1, qual2, qual3, q_cns;
    double rank_sum;
    double lk[10];
} orc_site;

void orc_posterior_site(orc_site* st, const uint32_t* rec, int n, const int32_t* next_q_sum, const orc_params* p,
                        const double* p_prior, const orc_snp* sn, const double* p_rank);

/* Build and call one window: sites [ws, ws+C) of chromosome c, from the
 * chromosome's valid reads idx[0..nidx) (pos >= 0, pos+len < chr_len, in
 * input order with non-decreasing positions).  'out' has
 * C+1 entries (the last one is the next site, whose q_sum the HET cap reads
 * through base codes 4..7).  Returns ORC_OK or an error. */
int orc_call_window(const orc_reads* r, const int64_t* idx, int64_t nidx, const orc_chr* c, uint32_t ws, uint32_t C,
                    const orc_params* p, const double* score, const double* p_prior, const double* p_rank,
                    orc_site* out, uint32_t* rec_out /* optional (C+1)*256 sorted records */,
                    int32_t* nrec_out);

/* .cns row text for one site; returns its length (no NUL counted). */
int orc_format_row(const char* chr_name, uint32_t pos1, const orc_site* s, const orc_site* next, char* out);

/* --------------------------------------------------------- -C codec pieces */
int orc_num_bits(uint64_t n_dict);
uint64_t orc_dict_encode_i32(const int32_t* in, uint64_t n, uint32_t* words, int32_t* dict, uint64_t* n_dict);
uint64_t orc_dict_encode_u32(const uint32_t* in, uint64_t n, uint32_t* words, uint32_t* dict, uint64_t* n_dict);
uint64_t orc_dict_encode_f64(const double* in, uint64_t n, uint32_t* words, double* dict, uint64_t* n_dict);
uint64_t orc_encode_2bit(const char* seq, uint64_t n, uint32_t* code);
void orc_decode_2bit(char* seq, uint64_t n, const uint32_t* code, uint64_t n_word);

/* -------------------------------------------------------------- pipeline */
typedef struct orc_run_stats {
    uint64_t sites;
    uint64_t windows;
    double seconds_call;
} orc_run_stats;

int orc_run(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in,
            const char* matrix_out, const orc_params* p, const char* out, int compress,
            int64_t max_sites, orc_run_stats* stats);
/* orc_run with the text windows called on nthreads threads (-C output stays
 * serial); seconds_call is then the wall time of the parallel section */
int orc_run_threads(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in,
                    const char* matrix_out, const orc_params* p, const char* out, int compress, int64_t max_sites,
                    int nthreads, orc_run_stats* stats);
/* orc_run_threads with -T regions: only windows a region touches are called */
int orc_run_opts(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in,
                 const char* matrix_out, const orc_params* p, const char* out, int compress, int64_t max_sites,
                 int nthreads, const char* regions, orc_run_stats* stats);
/* region masks (cpp/chromosome.cc:66-107): words of 64 positions, position p
 * at bit 63 - p % 64; either mask may be NULL */
int orc_region_set(uint64_t* site_mask, uint64_t* win_mask, uint32_t len, int start, int end);
int orc_region_in_win(const uint64_t* win_mask, uint32_t pos);
uint64_t orc_region_words(uint64_t n);
int orc_decompress(const char* prefix, const char* out);
/* the -C files (<prefix>.1_2 ... .15_17) of one window of given sites */
int orc_cwrite_sites(const char* prefix, const char* chr, uint32_t ws, uint32_t C, const orc_site* sites);
const char* orc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
