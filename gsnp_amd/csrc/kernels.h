// kernels.h -- device buffers and launch wrappers of the calling kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gsnp {

// Per-site counters, structure of arrays over a span of sites.
struct SiteArrays {
    uint32_t* rec = nullptr;  // [n][256] observation records, site-major
    int32_t* nrec = nullptr;  // records per site (== dep_uni)
    int32_t* depth = nullptr;
    int32_t* repeat = nullptr;
    int4* cu = nullptr;  // count_uni per base
    int4* qs = nullptr;  // q_sum per base
    int4* ca = nullptr;  // count_all per base
};

struct Tables {
    const double* score = nullptr;  // [256*256*4][10]
    const uint8_t* qadj = nullptr;  // [64][td][gd]
    int td = 1, gd = 1;
    const double* lprior = nullptr;  // [8][10]
    const double* prank = nullptr;   // 64*64*2048
    int read_len = 100;
    int q_min = 64;
    int q_max_min = 40;
    int monoploid = 0;
    int refine = 0;
    int rank_mode = 0;
};

struct RefView {
    const uint8_t* code = nullptr;  // chromosome codes
    uint32_t len = 0;
    const uint32_t* snp_pos = nullptr;
    const double* snp_lp = nullptr;  // [n_snp][10]
    uint32_t n_snp = 0;
    const uint32_t* snp_first = nullptr;  // [len/256 + 2]: first known SNP at or after each 256-site block
};

struct ReadView {
    const int32_t* pos = nullptr;
    const uint32_t* meta = nullptr;
    const uint8_t* seq = nullptr;
    uint32_t pitch = 0;
};

struct Results {
    uint8_t* base1 = nullptr;
    uint8_t* base2 = nullptr;
    uint8_t* type1 = nullptr;
    int32_t* q_cns = nullptr;
    double* rank = nullptr;
    double* lk = nullptr;  // [n][10]
};

// A het call that needs the rank-sum test (filled by the posterior kernel).
struct RankJob {
    uint32_t site;
    uint8_t type1, het_ok, pad0, pad1;
    int32_t cap;
    double dtp;  // type_prob[type1] - type_prob[type2]
};

// A site whose value depends on a host-libm transcendental (normal tail, or a
// log10 within the guard band of an integer): the host recomputes it.
struct FixJob {
    uint32_t site;
    int32_t n1, n2;
    uint8_t het_ok, normal, pad0, pad1;
    int32_t cap;
    double T1, T2, dtp, table_rank;
};

struct GText;

struct RowCtx {
    const char* name = nullptr;
    int name_len = 0;
    uint32_t pos_base = 0;         // 1-based position of local site 0
    const uint32_t* gidx = nullptr;  // per site: entry of gtext (valid where the fast %#g path fails)
    const GText* gtext = nullptr;
};

// error flags written by kernels
enum : uint32_t { kErrQadjRange = 1u, kErrRecordOverflow = 2u };

// Arguments of the fused sort + likelihood + posterior launch (batched form).
struct FusedArgs {
    Tables t;
    RefView ref;
    uint32_t ws = 0;
    SiteArrays a;
    const int4* head_qs = nullptr;
    Results out;
    RankJob* jobs = nullptr;
    uint32_t* job_count = nullptr;
    uint32_t* err = nullptr;
    int write_lk = 0;   // also store the ten log-likelihoods of each site
    int posterior = 1;  // call the sites (0: likelihoods only, for the per-window path)
};

// Launch shape of the fused stage: bucket classes below min_class run on
// min_class's network (the smallest bucket edge), and each class gets its
// waves per block (0: the most that fit 64 KB of LDS) and a grid cap.  The
// defaults are what gsnp_autotune picks on a 100 bp, 30x-mean contig of
// 8 Mb: sites of up to 8 records on the 16-wide network, four waves per
// block for that class and the wide ones, eight for the 32 and 64 classes.
struct FusedTune {
    uint32_t min_class = 1;
    uint32_t wpb[6] = {0, 4, 8, 8, 4, 4};
    uint32_t max_blocks[6] = {8192, 8192, 8192, 8192, 8192, 8192};
};

void launch_fill_u32(hipStream_t s, void* p, uint32_t n_words, uint32_t v);

// range: 2 u64 per 64-site block (read interval of the block)
void launch_site_build(hipStream_t s, const ReadView& r, uint64_t r0, uint64_t r1, uint32_t ws, uint32_t n_sites,
                       int64_t lo_min, int64_t hi_max, uint32_t carry, const SiteArrays& a, int4* head_qs,
                       uint64_t* range);
void launch_bucket(hipStream_t s, const int32_t* nrec, uint32_t n_sites, uint32_t* lists, uint32_t list_pitch,
                   uint32_t* counts, int all, unsigned long long* n_records, int min_class = 0);
void launch_sort(hipStream_t s, uint32_t* rec, const int32_t* nrec, const uint32_t* lists, uint32_t list_pitch,
                 const uint32_t* counts, uint32_t max_sites);
void launch_likelihood(hipStream_t s, const Tables& t, const SiteArrays& a, uint32_t n_sites, double* lk, uint32_t* err);
void launch_posterior(hipStream_t s, const Tables& t, const RefView& ref, uint32_t ws, const SiteArrays& a,
                      uint32_t n_sites, const int4* head_qs, Results& out, RankJob* jobs, uint32_t* job_count);
void launch_fused(hipStream_t s, const FusedArgs& f, const uint32_t* lists, uint32_t list_pitch, const uint32_t* counts,
                  uint32_t max_sites, const FusedTune& tune);
void launch_fused_one(hipStream_t s, const FusedArgs& f, const uint32_t* lists, uint32_t list_pitch,
                      const uint32_t* counts, uint32_t max_sites, uint32_t cls, uint32_t wpb, uint32_t max_blocks);
// normal_count: incremented once per site whose normal approximation the
// device finished
void launch_rank(hipStream_t s, const Tables& t, const SiteArrays& a, const RankJob* jobs, const uint32_t* job_count,
                 uint32_t max_jobs, Results& out, FixJob* fixes, uint32_t* fix_count, unsigned long long* normal_count);
void launch_apply_fix(hipStream_t s, const FixJob* fixes, const double* rank, const int32_t* qcns, uint32_t n,
                      Results& out);
// row lengths and block sums; rows that need the exact %#g routine are listed
// in gtext (gcount must be zero on entry) for launch_gtext
void launch_rowlen(hipStream_t s, const RefView& ref, uint32_t ws, const SiteArrays& a, const Results& res,
                   uint32_t n_sites, RowCtx rc, uint32_t* gidx, GText* gtext, uint32_t* gcount, uint32_t* rowlen,
                   uint64_t* block_sum);
// exact %#g text of the listed rows; their lengths are added to rowlen and block_sum
void launch_gtext(hipStream_t s, const SiteArrays& a, const Results& res, uint32_t n_sites, GText* gtext,
                  const uint32_t* gcount, uint32_t* rowlen, uint64_t* block_sum);
void launch_scan_blocks(hipStream_t s, uint64_t* block_sum, uint32_t n_blocks, uint64_t* total);
void launch_format(hipStream_t s, const RefView& ref, uint32_t ws, const SiteArrays& a, const Results& res,
                   uint32_t n_sites, RowCtx rc, const uint32_t* rowlen, const uint64_t* block_off, char* out);
void launch_recycle(hipStream_t s, const SiteArrays& a, uint32_t win, uint32_t L);

constexpr uint32_t kRowBlock = 256;

}  // namespace gsnp
