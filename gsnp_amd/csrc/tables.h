// tables.h -- host-built lookup tables for the calling kernels.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"

namespace gsnp {

// log10(0.5 p[q,c,a1,o] + 0.5 p[q,c,a2,o]) for every (q_adj, cycle, o) row and
// the 10 genotypes (kernel/likelihood_kernel.cu:137-161).  256*256*4 rows of
// kScoreStride doubles (10 used), so that every row is one 128-byte line.
constexpr uint32_t kScoreStride = 16;
std::vector<double> build_score_table(const double* p_matrix);

// Adjusted quality max(1, (int)(10^(log10 q + (t-1) pcr + g glob) + 0.5))
// (kernel/likelihood_kernel.cu:226-231), evaluated with the host libm for
// q in [0,64), t-1 in [0,td), g in [0,gd).  When a dependency coefficient is
// <= 0 the value stops changing along that axis, so lookups clamp t-1 to
// td-1 and g to gd-1 exactly.  Entries above 255 hold 0 (out of table).
struct QadjTable {
    int td = 1, gd = 1;
    std::vector<uint8_t> v;  // [q][t-1][g]
};
QadjTable build_qadj_table(double pcr, double glob);

// Genotype priors p_prior[ref(8)][a1<<2|a2] (cpp/prior.cc:3-43).
void prior_gen(const gsnp_params& p, double* p_prior /*128*/);

// log10 of the priors of the 10 genotypes for each of the 8 reference codes.
std::vector<double> log_prior_rows(const double* p_prior);

// log10 prior row of a known-SNP site (kernel/posterior_kernel.cu:575-633
// applied on top of the reference row, then log10).
void snp_log_prior(const gsnp_params& p, const double* p_prior, uint8_t ref_code, const gsnp_snp& s, double* out10);

// Wilcoxon rank-sum null tables (cpp/rank_sum.cc:3-42), 64*64*2048 doubles.
std::vector<double> rank_table();

}  // namespace gsnp
