// tables.cpp -- host-built lookup tables for the calling kernels.  Compiled
// with -ffp-contract=off: every value here feeds an exact comparison or an
// integer truncation on the device, so it must be the plain IEEE result of
// the host libm.
#include "tables.h"

#include <cmath>
#include <cstring>

namespace gsnp {

std::vector<double> build_score_table(const double* pm)
{
    std::vector<double> s((size_t)256 * 256 * 4 * kScoreStride, 0.0);
    for (uint32_t q = 0; q < 256; q++) {
        for (uint32_t c = 0; c < 256; c++) {
            const double* row = pm + ((size_t)q << 12) + ((size_t)c << 4);
            for (uint32_t o = 0; o < 4; o++) {
                double* out = &s[((size_t)((q << 10) | (c << 2) | o)) * kScoreStride];
                int g = 0;
                for (uint32_t a1 = 0; a1 < 4; a1++)
                    for (uint32_t a2 = a1; a2 < 4; a2++) {
                        double half1 = 0.5 * row[(a1 << 2) | o];
                        double half2 = 0.5 * row[(a2 << 2) | o];
                        out[g++] = std::log10(half1 + half2);
                    }
            }
        }
    }
    // Row 0 (adjusted quality 0, cycle 0, base A) is never a record's row:
    // adjusted qualities are at least 1.  It holds +0.0, the value the fused
    // kernel adds for the padding slots of a site.
    for (uint32_t g = 0; g < kScoreStride; g++) s[g] = 0.0;
    return s;
}

static int qadj_value(int q, int tm1, int g, double lq, double pcr, double glob)
{
    (void)q;
    double x = lq + tm1 * pcr;
    x = x + g * glob;
    int v = (int)(std::pow(10.0, x) + 0.5);
    return v < 1 ? 1 : v;
}

QadjTable build_qadj_table(double pcr, double glob)
{
    QadjTable t;
    double lq[64];
    for (int q = 0; q < 64; q++) lq[q] = std::log10((double)q);
    // extent along an axis: 1 if the coefficient is 0, the first index where
    // every q has reached the floor 1 if it is negative, the full 256 otherwise
    auto extent = [&](bool along_t, double coef) {
        if (coef == 0.0) return 1;
        if (!(coef < 0.0)) return 256;
        int need = 1;
        for (int q = 1; q < 64; q++) {
            int i = 0;
            while (i < 255 && (along_t ? qadj_value(q, i, 0, lq[q], pcr, glob) : qadj_value(q, 0, i, lq[q], pcr, glob)) > 1)
                i++;
            if (i + 1 > need) need = i + 1;
        }
        return need;
    };
    t.td = extent(true, pcr);
    t.gd = extent(false, glob);
    t.v.assign((size_t)64 * t.td * t.gd, 0);
    for (int q = 0; q < 64; q++)
        for (int a = 0; a < t.td; a++)
            for (int g = 0; g < t.gd; g++) {
                int v = qadj_value(q, a, g, lq[q], pcr, glob);
                t.v[((size_t)q * t.td + a) * t.gd + g] = (uint8_t)(v > 255 ? 0 : v);
            }
    return t;
}

void prior_gen(const gsnp_params& p, double* pp)
{
    for (int i = 0; i < 128; i++) pp[i] = 1.0;
    for (int r = 0; r < 4; r++)
        for (int a1 = 0; a1 < 4; a1++)
            for (int a2 = a1; a2 < 4; a2++) {
                double v;
                bool has1 = a1 == r, has2 = a2 == r;
                if (has1 && has2)
                    v = 1;  // ref hom
                else if (has1 || has2)
                    v = p.het_novel_r;  // ref het
                else if (a1 == a2)
                    v = p.althom_novel_r;  // alt hom
                else
                    v = p.het_novel_r * p.althom_novel_r;  // two-alt het
                if (p.transition_dominant && ((a1 ^ r) == 3 || (a2 ^ r) == 3)) v *= 4;
                pp[(r << 4) | (a1 << 2) | a2] = v;
            }
    for (int r = 4; r < 8; r++)
        for (int a1 = 0; a1 < 4; a1++)
            for (int a2 = a1; a2 < 4; a2++) pp[(r << 4) | (a1 << 2) | a2] = (a1 == a2 ? 1 : (2 * p.het_novel_r)) * 0.25 * 0.25;
}

static const int kGeno[10] = {0, 1, 2, 3, 5, 6, 7, 10, 11, 15};

std::vector<double> log_prior_rows(const double* pp)
{
    std::vector<double> lp(8 * 10);
    for (int r = 0; r < 8; r++)
        for (int g = 0; g < 10; g++) lp[r * 10 + g] = std::log10(pp[(r << 4) | kGeno[g]]);
    return lp;
}

void snp_log_prior(const gsnp_params& p, const double* pp, uint8_t ref, const gsnp_snp& s, double* out10)
{
    double row[16];
    std::memcpy(row, pp + ((ref & 7) << 4), sizeof row);
    if (!s.indel) {
        int t = ref & 3;
        for (int a1 = 0; a1 < 4; a1++)
            for (int a2 = a1; a2 < 4; a2++) {
                if (!(s.freq[a1] > 0 && s.freq[a2] > 0)) continue;
                double v;
                if (s.hapmap) {
                    v = (a1 == a2 ? 1 : (2 * p.het_val_r)) * s.freq[a1] * s.freq[a2];
                } else if (a1 == a2 && a1 == t) {
                    v = 1;
                } else if (a1 == t || a2 == t) {
                    v = s.validated ? p.het_val_r : p.het_unval_r;
                } else if (a1 == a2) {
                    v = s.validated ? p.althom_val_r : p.althom_unval_r;
                } else {
                    v = s.validated ? p.het_val_r : p.het_unval_r;
                }
                row[(a1 << 2) | a2] = v;
            }
    }
    for (int g = 0; g < 10; g++) out10[g] = std::log10(row[kGeno[g]]);
}

std::vector<double> rank_table()
{
    const size_t kSize = (size_t)64 * 64 * 2048;
    std::vector<double> pr(kSize, 1.0);
    std::vector<uint64_t> cnt(kSize, 0);
    double fact[64];
    fact[0] = 1.0;
    for (int i = 1; i < 64; i++) fact[i] = fact[i - 1] * i;
    auto at = [](uint64_t N, uint64_t n1, uint64_t T) { return (N << 17) | (n1 << 11) | T; };
    cnt[0] = 1;
    for (uint64_t N = 1; N < 64; N++) {
        for (uint64_t n1 = 0; n1 <= N; n1++) {
            const uint64_t lo = (1 + n1) * n1 / 2, hi = (2 * N - n1 + 1) * n1 / 2;
            // number of n1-subsets of 1..N with sum T: either N is out, or N is in
            for (uint64_t T = lo; T <= hi; T++) {
                uint64_t c = cnt[at(N - 1, n1, T)];
                if (T >= N && n1 > 0) c += cnt[at(N - 1, n1 - 1, T - N)];
                cnt[at(N, n1, T)] = c;
                pr[at(N, n1, T)] = c / (fact[N] / (fact[n1] * fact[N - n1]));
            }
            // two-sided: smaller of the running left tail and the remaining right tail
            double left = 0.0, right = 1.0;
            for (uint64_t T = lo; T <= hi; T++) {
                right = 1.0 - left;
                left += pr[at(N, n1, T)];
                pr[at(N, n1, T)] = left < right ? left : right;
            }
        }
    }
    return pr;
}

}  // namespace gsnp
