// engine.hip -- context, C ABI and the per-window driver of the calling path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "cwriter.h"
#include "engine.h"
#include "gsnp_common.h"
#include "inputs.h"
#include "kernels.h"
#include "tables.h"

namespace gsnp {

static_assert(kScoreStride == kScoreRowStride, "host and device score-row strides differ");

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

hipEvent_t StageTimer::take()
{
    if (!pool.empty()) {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void StageTimer::begin(int st, hipStream_t s)
{
    if (!on) return;
    open[st] = take();
    if (open[st]) (void)hipEventRecord(open[st], s);
}

void StageTimer::end(int st, hipStream_t s)
{
    if (!on || !open[st]) return;
    hipEvent_t e = take();
    if (!e) return;
    (void)hipEventRecord(e, s);
    pending[st].push_back({open[st], e});
    open[st] = nullptr;
}

int StageTimer::collect()
{
    for (int st = 0; st < kNumStages; st++) {
        for (auto& pr : pending[st]) {
            float t = 0.f;
            HIPCHK(hipEventSynchronize(pr.second));
            HIPCHK(hipEventElapsedTime(&t, pr.first, pr.second));
            ms[st] += t;
            count[st] += 1;
            pool.push_back(pr.first);
            pool.push_back(pr.second);
        }
        pending[st].clear();
    }
    return GSNP_OK;
}

StageTimer::~StageTimer()
{
    for (int st = 0; st < kNumStages; st++)
        for (auto& pr : pending[st]) pool.push_back(pr.first), pool.push_back(pr.second);
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

int Ctx::alloc_window(uint32_t n_sites)
{
    if (n_sites <= win_cap) return GSNP_OK;
    HIPCHK(hipStreamSynchronize(stream));  // nothing queued may still use the old buffers
    buf_gen++;
    sites.reset();
    win_cap = 0;
    const size_t n = n_sites;
    HIPCHK(sites.rec.alloc(n * kSlots));
    HIPCHK(sites.nrec.alloc(n));
    HIPCHK(sites.depth.alloc(n));
    HIPCHK(sites.repeat.alloc(n));
    HIPCHK(sites.cu.alloc(n));
    HIPCHK(sites.qs.alloc(n));
    HIPCHK(sites.ca.alloc(n));
    HIPCHK(lk.alloc(n * 10));
    HIPCHK(base1.alloc(n));
    HIPCHK(base2.alloc(n));
    HIPCHK(type1.alloc(n));
    HIPCHK(qcns.alloc(n));
    HIPCHK(rank.alloc(n));
    HIPCHK(lists.alloc(n * 6));
    HIPCHK(jobs.alloc(n));
    HIPCHK(fixes.alloc(n));
    HIPCHK(rowlen.alloc(n));
    HIPCHK(blocksum.alloc(n / kRowBlock + 2));
    HIPCHK(ranges.alloc(2 * (n / 64 + 2)));
    HIPCHK(gidx.alloc(n));
    HIPCHK(gtext.alloc(n));
    HIPCHK(hipMemsetAsync(sites.nrec.p, 0, n * 4, stream));
    HIPCHK(hipMemsetAsync(sites.depth.p, 0, n * 4, stream));
    HIPCHK(hipMemsetAsync(sites.repeat.p, 0, n * 4, stream));
    HIPCHK(hipMemsetAsync(sites.cu.p, 0, n * 16, stream));
    HIPCHK(hipMemsetAsync(sites.qs.p, 0, n * 16, stream));
    HIPCHK(hipMemsetAsync(sites.ca.p, 0, n * 16, stream));
    win_cap = n_sites;
    return GSNP_OK;
}

SiteArrays Ctx::site_view(uint32_t offset)
{
    SiteArrays a;
    a.rec = sites.rec.p + (size_t)offset * kSlots;
    a.nrec = sites.nrec.p + offset;
    a.depth = sites.depth.p + offset;
    a.repeat = sites.repeat.p + offset;
    a.cu = sites.cu.p + offset;
    a.qs = sites.qs.p + offset;
    a.ca = sites.ca.p + offset;
    return a;
}

Results Ctx::result_view(uint32_t offset)
{
    Results r;
    r.base1 = base1.p + offset;
    r.base2 = base2.p + offset;
    r.type1 = type1.p + offset;
    r.q_cns = qcns.p + offset;
    r.rank = rank.p + offset;
    r.lk = lk.p + (size_t)offset * 10;
    return r;
}

Tables Ctx::tables() const
{
    Tables t;
    t.score = score.p;
    t.qadj = qadj.p;
    t.td = td;
    t.gd = gd;
    t.lprior = lprior.p;
    t.prank = prank.p;
    t.read_len = p.read_length;
    t.q_min = p.q_min;
    t.q_max_min = p.q_max - p.q_min;
    t.monoploid = p.is_monoploid;
    t.refine = p.refine_mode;
    t.rank_mode = p.rank_sum_mode;
    return t;
}

RefView Ctx::ref_view() const
{
    RefView r;
    r.code = ref.p;
    r.len = chr_len;
    r.snp_pos = snp_pos.p;
    r.snp_first = snp_first.p;
    r.snp_lp = snp_lp.p;
    r.n_snp = n_snp;
    return r;
}

ReadView Ctx::read_view() const
{
    ReadView v;
    v.pos = rpos.p;
    v.meta = rmeta.p;
    v.seq = rseq.p;
    v.pitch = pitch ? pitch : 1u;
    return v;
}

// (int)x as x86-64 cvttsd2si
static int trunc_x86(double x)
{
    if (!(x > -2147483649.0 && x < 2147483648.0)) return (int)0x80000000;
    return (int)x;
}

// cpp/normal_dis.cc:3-24, cpp/rank_sum.cc:44-49 with the host libm
static double host_normal_value(double z)
{
    if (z > 6.0 || z < -6.0) return 0.0;
    const double b1 = 0.31938153, b2 = -0.356563782, b3 = 1.781477937, b4 = -1.821255978, b5 = 1.330274429;
    const double pp = 0.2316419, c2 = 0.39894228;
    double a = std::fabs(z);
    double t = 1.0 / (1.0 + a * pp);
    double b = c2 * std::exp((-z) * (z / 2.0));
    double n = ((((b5 * t + b4) * t + b3) * t + b2) * t + b1) * t;
    n = 1.0 - b * n;
    if (z < 0.0) n = 1.0 - n;
    return n > 0.5 ? 1 - n : n;
}

static double host_normal_test(int n1, int n2, double T1, double T2)
{
    double u1 = (T1 - n1 * (n1 + n2 + 1) / 2) / std::sqrt(n1 * n2 * (n1 + n2 + 1) / (double)12);
    double u2 = (T2 - n2 * (n1 + n2 + 1) / 2) / std::sqrt(n1 * n2 * (n1 + n2 + 1) / (double)12);
    return host_normal_value(std::fabs(u1) > std::fabs(u2) ? u1 : u2);
}

static double host_table_test(const double* p, int n1, int n2, double T1, double T2)
{
    if (n1 <= n2)
        return p[(n1 + n2) << 17 | n1 << 11 | (int)(T1)] +
               (T1 - (int)T1) * (p[(n1 + n2) << 16 | n1 << 11 | (int)(T1 + 1)] - p[(n1 + n2) << 17 | n1 << 11 | (int)(T1)]);
    return p[(n1 + n2) << 17 | n2 << 11 | (int)(T2)] +
           (T2 - (int)T2) * (p[(n1 + n2) << 16 | n2 << 11 | (int)(T2 + 1)] - p[(n1 + n2) << 17 | n2 << 11 | (int)(T2)]);
}

// Sites whose rank-sum value or consensus quality depends on a host-libm
// transcendental are finished here with the host libm, then patched back.
int Ctx::host_fixups()
{
    uint32_t nfix = 0;
    HIPCHK(hipMemcpyAsync(&nfix, fix_count.p, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    last_fixups = nfix;
    if (!nfix) return GSNP_OK;
    std::vector<FixJob> fx(nfix);
    HIPCHK(hipMemcpyAsync(fx.data(), fixes.p, sizeof(FixJob) * nfix, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<double> r(nfix);
    std::vector<int32_t> q(nfix);
    fixups_total += nfix;
    for (uint32_t i = 0; i < nfix; i++) {
        const FixJob& f = fx[i];
        fixups_normal += f.normal ? 1 : 0;
        double rk = f.normal ? host_normal_test(f.n1, f.n2, f.T1, f.T2)
                             : host_table_test(h_prank.data(), f.n1, f.n2, f.T1, f.T2);
        int qc = rk == 0.0 ? 0 : trunc_x86(10 * f.dtp + 10 * std::log10(rk));
        if (f.het_ok) {
            if (qc > f.cap) qc = f.cap;
        } else {
            qc = 0;
        }
        if (qc > 99) qc = 99;
        if (qc < 0) qc = 0;
        r[i] = rk;
        q[i] = qc;
    }
    DevBuf<double> dr;
    DevBuf<int32_t> dq;
    HIPCHK(dr.alloc(nfix));
    HIPCHK(dq.alloc(nfix));
    HIPCHK(hipMemcpyAsync(dr.p, r.data(), 8 * nfix, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(dq.p, q.data(), 4 * nfix, hipMemcpyHostToDevice, stream));
    Results res = result_view(fix_offset);
    launch_apply_fix(stream, fixes.p, dr.p, dq.p, nfix, res);
    HIPCHK(hipStreamSynchronize(stream));
    return GSNP_OK;
}

int Ctx::check_err()
{
    uint32_t e = 0;
    HIPCHK(hipMemcpyAsync(&e, err.p, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (e & kErrQadjRange) return fail(GSNP_E_RANGE, "adjusted quality exceeds the score table");
    return GSNP_OK;
}

void Ctx::swap_chrom(ChromState& s)
{
    std::swap(chrom_id, s.chrom_id);
    std::swap(chr_name, s.chr_name);
    std::swap(chr_len, s.chr_len);
    ref.swap(s.ref);
    snp_pos.swap(s.snp_pos);
    snp_first.swap(s.snp_first);
    snp_lp.swap(s.snp_lp);
    std::swap(n_snp, s.n_snp);
    d_name.swap(s.d_name);
    std::swap(n_reads, s.n_reads);
    std::swap(pitch, s.pitch);
    rpos.swap(s.rpos);
    rmeta.swap(s.rmeta);
    rseq.swap(s.rseq);
    h_rpos.swap(s.h_rpos);
    h_snps.swap(s.h_snps);
    h_snp_ref.swap(s.h_snp_ref);
    built = false;
    adj_n = 0;  // a prepared adjoint span belongs to the chromosome it was built on
}

static uint64_t next_chrom_id()
{
    static std::atomic<uint64_t> id{0};
    return ++id;
}

// The score table, from a calibration matrix (256*256*16 doubles); every other
// table is built at creation.
int Ctx::set_matrix(const double* p_matrix)
{
    std::vector<double> sc = build_score_table(p_matrix);
    HIPCHK(hipStreamSynchronize(stream));
    if (score.n != sc.size()) HIPCHK(score.alloc(sc.size()));
    HIPCHK(hipMemcpyAsync(score.p, sc.data(), sc.size() * 8, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return GSNP_OK;
}

// A context; with p_matrix null its score table is left for set_matrix (a job
// that counts its calibration on the device creates the context first).
int create_ctx(const gsnp_params* p, const double* p_matrix, gsnp_ctx** out)
{
    if (!p || !out) return fail(GSNP_E_ARG, "null argument");
    if (p->read_length < 1 || p->read_length > 256) return fail(GSNP_E_ARG, "read length outside [1,256]");
    if (p->q_min < 1 || p->q_max < p->q_min || p->q_max > 126 || p->q_max - p->q_min > 62)
        return fail(GSNP_E_ARG, "quality characters out of range");
    std::unique_ptr<gsnp_ctx> c(new gsnp_ctx());
    c->p = *p;
    HIPCHK(hipGetDevice(&c->device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = c->stream;
    // adjusted qualities, priors, rank table (the score table: set_matrix)
    QadjTable qt = build_qadj_table(p->pcr_dependency, p->global_dependency);
    c->td = qt.td;
    c->gd = qt.gd;
    c->h_pprior.resize(128);
    prior_gen(*p, c->h_pprior.data());
    std::vector<double> lp = log_prior_rows(c->h_pprior.data());
    c->h_prank = rank_table();
    HIPCHK(c->qadj.alloc(qt.v.size()));
    HIPCHK(c->lprior.alloc(lp.size()));
    HIPCHK(c->prank.alloc(c->h_prank.size()));
    HIPCHK(c->err.alloc(1));
    HIPCHK(c->counts.alloc(8));
    HIPCHK(c->job_count.alloc(1));
    HIPCHK(c->fix_count.alloc(1));
    HIPCHK(c->total.alloc(1));
    HIPCHK(c->stat.alloc(4));
    HIPCHK(hipMemsetAsync(c->stat.p, 0, 32, c->stream));
    HIPCHK(hipMemcpyAsync(c->qadj.p, qt.v.data(), qt.v.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->lprior.p, lp.data(), lp.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->prank.p, c->h_prank.data(), c->h_prank.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->err.p, 0, 4, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (p_matrix) {
        int rc = c->set_matrix(p_matrix);
        if (rc) return rc;
    }
    *out = c.release();
    return GSNP_OK;
}

}  // namespace gsnp

using namespace gsnp;

extern "C" {

void gsnp_params_default(gsnp_params* p)
{
    std::memset(p, 0, sizeof *p);
    p->q_min = 64;
    p->q_max = 64 + 40;
    p->read_length = 45;
    p->althom_novel_r = 0.0005;
    p->het_novel_r = 0.0010;
    p->althom_val_r = 0.05;
    p->het_val_r = 0.10;
    p->althom_unval_r = 0.01;
    p->het_unval_r = 0.02;
    p->global_dependency = std::log10(0.9);
    p->pcr_dependency = std::log10(0.5);
}

const char* gsnp_last_error(void) { return g_err.c_str(); }

int gsnp_device_init(int device)
{
    HIPCHK(hipSetDevice(device));
    uint32_t* scratch = nullptr;  // per-device address of the no-allocation entry point's array
    return noalloc_scratch(device, &scratch);
}

int gsnp_create(const gsnp_params* p, const double* p_matrix, gsnp_ctx** out)
{
    if (!p || !p_matrix || !out) return fail(GSNP_E_ARG, "null argument");
    return create_ctx(p, p_matrix, out);
}

int gsnp_set_matrix(gsnp_ctx* c, const double* p_matrix)
{
    if (!c || !p_matrix) return fail(GSNP_E_ARG, "null argument");
    return c->set_matrix(p_matrix);
}

void gsnp_destroy(gsnp_ctx* c)
{
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    hipStream_t s = c->own_stream;
    delete c;
    if (s) hipStreamDestroy(s);
}

int gsnp_set_stream(gsnp_ctx* c, void* stream)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    hipStream_t next = stream ? static_cast<hipStream_t>(stream) : c->own_stream;
    if (next == c->stream) return GSNP_OK;
    // the work queued so far ends before the next call's work starts
    HIPCHK(hipStreamSynchronize(c->stream));
    c->stream = next;
    return GSNP_OK;
}

int gsnp_set_chromosome(gsnp_ctx* c, const char* name, const uint8_t* codes, uint32_t len, const uint32_t* snp_pos,
                        const gsnp_snp* snps, uint32_t n_snp)
{
    if (!c || !name || (!codes && len)) return fail(GSNP_E_ARG, "null argument");
    if (n_snp && (!snp_pos || !snps)) return fail(GSNP_E_ARG, "null known-SNP arrays");
    // everything is checked before the context changes: a rejected call leaves
    // the previous chromosome in place
    for (uint32_t i = 0; i < n_snp; i++)
        if (snp_pos[i] >= len || (i && snp_pos[i] <= snp_pos[i - 1]))
            return fail(GSNP_E_ARG, "known-SNP positions must be sorted, unique and inside the chromosome");
    HIPCHK(hipStreamSynchronize(c->stream));
    // until every upload below has succeeded the context holds no chromosome
    // (length 0, no known SNPs, no reads): a failure leaves nothing stale to call
    c->chr_len = 0;
    c->n_snp = 0;
    c->n_reads = 0;
    c->buf_gen++;
    c->chrom_id = next_chrom_id();
    c->chr_name = name;
    HIPCHK(c->ref.alloc((size_t)len + GSNP_WIN + 512));
    HIPCHK(hipMemsetAsync(c->ref.p, 4, (size_t)len + GSNP_WIN + 512, c->stream));  // beyond the end reads as N
    if (len) HIPCHK(hipMemcpyAsync(c->ref.p, codes, len, hipMemcpyHostToDevice, c->stream));
    c->h_snps.assign(snps, snps + n_snp);
    c->h_snp_ref.resize(n_snp);
    for (uint32_t i = 0; i < n_snp; i++) c->h_snp_ref[i] = snp_pos[i] < len ? codes[snp_pos[i]] : 0;
    if (n_snp) {
        std::vector<double> lp((size_t)n_snp * 10);
        for (uint32_t i = 0; i < n_snp; i++)
            snp_log_prior(c->p, c->h_pprior.data(), codes[snp_pos[i]], snps[i], &lp[(size_t)i * 10]);
        HIPCHK(c->snp_pos.alloc(n_snp));
        HIPCHK(c->snp_lp.alloc(lp.size()));
        HIPCHK(hipMemcpyAsync(c->snp_pos.p, snp_pos, 4ull * n_snp, hipMemcpyHostToDevice, c->stream));
        // first known SNP at or after each 256-site block (lookup index of the posterior)
        const size_t nb = (size_t)len / 256 + 2;
        std::vector<uint32_t> first(nb);
        uint32_t j = 0;
        for (size_t b = 0; b < nb; b++) {
            while (j < n_snp && snp_pos[j] < (uint64_t)b * 256) j++;
            first[b] = j;
        }
        HIPCHK(c->snp_first.alloc(nb));
        HIPCHK(hipMemcpyAsync(c->snp_first.p, first.data(), 4 * nb, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->snp_lp.p, lp.data(), 8 * lp.size(), hipMemcpyHostToDevice, c->stream));
    }
    const size_t nl = std::strlen(name);
    HIPCHK(c->d_name.alloc(nl + 1));
    HIPCHK(hipMemcpyAsync(c->d_name.p, name, nl + 1, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->chr_len = len;
    c->n_snp = n_snp;
    return GSNP_OK;
}

static bool is_acgt(char ch)
{
    switch (ch) {
    case 'A': case 'C': case 'G': case 'T': case 'a': case 'c': case 'g': case 't':
        return true;
    default:
        return false;
    }
}

int gsnp_load_reads(gsnp_ctx* c, const gsnp_read_batch* b)
{
    if (!c || !b) return fail(GSNP_E_ARG, "null argument");
    const uint32_t L = (uint32_t)c->p.read_length;
    if (b->n && (b->pitch < 1 || !b->pos || !b->len_hit || !b->strand || !b->seq || !b->qual))
        return fail(GSNP_E_ARG, "incomplete read batch");
    // validate and pack: byte = (q << 2) | base, 0 when the base is not observed;
    // ranges of reads on threads, the error of the lowest read index reported
    std::vector<uint32_t> meta(b->n);
    std::unique_ptr<uint8_t[]> seq(new uint8_t[(size_t)b->n * L + 1]);
    struct PackErr {
        uint64_t at = UINT64_MAX;
        int code = 0;
        std::string msg;
    };
    const uint32_t q_min = (uint32_t)c->p.q_min;
    const uint64_t chr_len = c->chr_len;
    auto pack = [&](uint64_t i0, uint64_t i1, PackErr& e) {
        int32_t last = i0 ? b->pos[i0 - 1] : 0;  // what the reads before i0 left, when they pass
        for (uint64_t i = i0; i < i1; i++) {
            const uint32_t lh = (uint32_t)b->len_hit[i];
            const uint32_t len = lh >> 16, hit = lh & 0xffff;
            const int32_t pos = b->pos[i];
            if (len < 1 || len > L || len > b->pitch) {
                e = {i, GSNP_E_RANGE, "read length outside [1, read_length]"};
                return;
            }
            if (pos < 0 || (uint64_t)pos + len >= chr_len) {
                e = {i, GSNP_E_RANGE, "read outside the chromosome (filter pos >= 0 and pos+len < length first)"};
                return;
            }
            if (pos < last) {
                e = {i, GSNP_E_UNSORTED, "Errors in sorting:" + std::to_string(pos) + "<" + std::to_string(last)};
                return;
            }
            last = pos;
            const bool rev = b->strand[i] != '+';
            meta[i] = hit | (len << 16) | (rev ? 0x80000000u : 0u);
            const char* sq = b->seq + i * b->pitch;
            const char* q = b->qual + i * b->pitch;
            uint8_t* o = &seq[i * L];
            std::memset(o, 0, L);
            for (uint32_t k = 0; k < len; k++) {
                const uint32_t qc = (unsigned char)q[k];
                if (!is_acgt(sq[k]) || qc < q_min) continue;
                uint32_t qq = qc - q_min;
                if (hit == 1) {
                    if (qq < 1 || qq > 63) {
                        e = {i, GSNP_E_QUAL, "quality outside [1,63] on a unique-hit base"};
                        return;
                    }
                } else {
                    qq = 1;  // repeats only count the base; any non-zero quality marks it observed
                }
                o[k] = (uint8_t)((qq << 2) | (((unsigned)sq[k] >> 1) & 3));
            }
        }
    };
    {
        const uint64_t hw = std::max(1u, std::thread::hardware_concurrency());
        const uint64_t T = std::max<uint64_t>(1, std::min<uint64_t>({hw, 16, b->n / 65536 + 1}));
        std::vector<PackErr> errs(T);
        std::vector<std::thread> th;
        std::vector<char> started(T, 0);  // a range whose thread cannot be started is packed here afterwards
        for (uint64_t t = 1; t < T; t++) {
            try {
                th.emplace_back(pack, b->n * t / T, b->n * (t + 1) / T, std::ref(errs[t]));
                started[t] = 1;
            } catch (const std::system_error&) {
            }
        }
        pack(0, b->n / T, errs[0]);
        for (auto& x : th) x.join();
        for (uint64_t t = 1; t < T; t++)
            if (!started[t]) pack(b->n * t / T, b->n * (t + 1) / T, errs[t]);
        for (const PackErr& e : errs)  // ranges are in read order: the first error found is the lowest
            if (e.code) return fail(e.code, e.msg);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->n_reads = 0;  // set once the uploads below have succeeded
    c->pitch = L;
    c->buf_gen++;
    c->chrom_id = next_chrom_id();
    c->h_rpos.assign(b->pos, b->pos + b->n);
    HIPCHK(c->rpos.alloc(b->n + 1));
    HIPCHK(c->rmeta.alloc(b->n + 1));
    const size_t nseq = (size_t)b->n * L;
    HIPCHK(c->rseq.alloc(nseq + 64));
    if (b->n) {
        HIPCHK(hipMemcpyAsync(c->rpos.p, b->pos, 4 * b->n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->rmeta.p, meta.data(), 4 * b->n, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->rseq.p, seq.get(), nseq, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->n_reads = b->n;
    return GSNP_OK;
}

int gsnp_window_init(gsnp_ctx* c, uint32_t start)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    if (!c->ref.p || !c->d_name.p) return fail(GSNP_E_ARG, "no chromosome set");
    const uint32_t L = (uint32_t)c->p.read_length;
    int rc = c->alloc_window(GSNP_WIN + L + 1);
    if (rc) return rc;
    c->ws = start;
    // only the carried head [0, L) is read before being written
    HIPCHK(hipMemsetAsync(c->sites.nrec.p, 0, 4 * (L + 1), c->stream));
    HIPCHK(hipMemsetAsync(c->sites.depth.p, 0, 4 * (L + 1), c->stream));
    HIPCHK(hipMemsetAsync(c->sites.repeat.p, 0, 4 * (L + 1), c->stream));
    HIPCHK(hipMemsetAsync(c->sites.cu.p, 0, 16 * (L + 1), c->stream));
    HIPCHK(hipMemsetAsync(c->sites.qs.p, 0, 16 * (L + 1), c->stream));
    HIPCHK(hipMemsetAsync(c->sites.ca.p, 0, 16 * (L + 1), c->stream));
    c->built = false;
    return GSNP_OK;
}

int gsnp_window_update(gsnp_ctx* c, uint64_t r0, uint64_t r1, uint32_t* max_records)
{
    if (!c || !c->win_cap) return fail(GSNP_E_ARG, "window not initialised");
    if (r0 > r1 || r1 > c->n_reads) return fail(GSNP_E_ARG, "read range outside the loaded batch");
    const uint32_t L = (uint32_t)c->p.read_length;
    c->timer.begin(kStSiteBuild, c->stream);
    launch_site_build(c->stream, c->read_view(), r0, r1, c->ws, GSNP_WIN + L, (int64_t)c->ws,
                      (int64_t)c->ws + GSNP_WIN - 1, L, c->site_view(0), nullptr, c->ranges.p);
    c->timer.end(kStSiteBuild, c->stream);
    HIPCHK(hipGetLastError());
    c->built = true;
    if (max_records) {
        std::vector<int32_t> n(GSNP_WIN);
        HIPCHK(hipMemcpyAsync(n.data(), c->sites.nrec.p, 4ull * GSNP_WIN, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        *max_records = (uint32_t)*std::max_element(n.begin(), n.end());
    }
    return GSNP_OK;
}

int gsnp_window_likelihood(gsnp_ctx* c, uint32_t C)
{
    if (!c || !c->win_cap) return fail(GSNP_E_ARG, "window not initialised");
    if (C > GSNP_WIN) return fail(GSNP_E_ARG, "call length above the window size");
    if (!c->built) {  // a window without reads still carries the previous tail
        int rc = gsnp_window_update(c, 0, 0, nullptr);
        if (rc) return rc;
    }
    SiteArrays a = c->site_view(0);
    // the fused kernel without its posterior: every site of the window is
    // bucketed, sorted in registers (the records in memory keep their order:
    // nothing downstream depends on it) and its ten log-likelihoods stored
    c->timer.begin(kStSort, c->stream);
    launch_bucket(c->stream, a.nrec, C, c->lists.p, c->win_cap, c->counts.p, 1, c->stat.p, (int)c->tune.min_class);
    c->timer.end(kStSort, c->stream);
    FusedArgs f = c->fused_args(c->ws);
    f.head_qs = nullptr;
    f.write_lk = 1;
    f.posterior = 0;
    c->timer.begin(kStLikelihood, c->stream);
    launch_fused(c->stream, f, c->lists.p, c->win_cap, c->counts.p, C, c->tune);
    c->timer.end(kStLikelihood, c->stream);
    HIPCHK(hipGetLastError());
    return GSNP_OK;
}

int gsnp_window_posterior(gsnp_ctx* c, uint32_t C)
{
    if (!c || !c->win_cap) return fail(GSNP_E_ARG, "window not initialised");
    if (C > GSNP_WIN) return fail(GSNP_E_ARG, "call length above the window size");
    SiteArrays a = c->site_view(0);
    Results res = c->result_view(0);
    HIPCHK(hipMemsetAsync(c->job_count.p, 0, 4, c->stream));
    HIPCHK(hipMemsetAsync(c->fix_count.p, 0, 4, c->stream));
    c->timer.begin(kStPosterior, c->stream);
    launch_posterior(c->stream, c->tables(), c->ref_view(), c->ws, a, C, nullptr, res, c->jobs.p, c->job_count.p);
    c->timer.end(kStPosterior, c->stream);
    if (c->p.rank_sum_mode) {
        c->timer.begin(kStRank, c->stream);
        launch_rank(c->stream, c->tables(), a, c->jobs.p, c->job_count.p, C, res, c->fixes.p, c->fix_count.p,
                    c->stat.p + 1);
        c->timer.end(kStRank, c->stream);
        HIPCHK(hipGetLastError());
        c->fix_offset = 0;
        int rc = c->host_fixups();
        if (rc) return rc;
    }
    HIPCHK(hipGetLastError());
    return c->check_err();
}

int gsnp_window_text(gsnp_ctx* c, uint32_t C, uint64_t* nbytes)
{
    if (!c || !c->win_cap) return fail(GSNP_E_ARG, "window not initialised");
    return c->format_text(0, c->ws, C, c->ws + 1, nbytes);
}

int gsnp_fetch_text(gsnp_ctx* c, char* dst, uint64_t cap, uint64_t* nbytes)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    if (nbytes) *nbytes = c->text_len;
    if (!dst) return GSNP_OK;
    if (cap < c->text_len) return fail(GSNP_E_ARG, "destination smaller than the formatted text");
    if (c->text_len) HIPCHK(hipMemcpyAsync(dst, c->text.p, c->text_len, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GSNP_OK;
}

int gsnp_window_results(gsnp_ctx* c, uint32_t C, const gsnp_window_out* o)
{
    if (!c || !o) return fail(GSNP_E_ARG, "null argument");
    auto cp = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (!dst || !bytes) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
    };
    HIPCHK(cp(o->base1, c->base1.p, C));
    HIPCHK(cp(o->base2, c->base2.p, C));
    HIPCHK(cp(o->type1, c->type1.p, C));
    HIPCHK(cp(o->q_cns, c->qcns.p, 4ull * C));
    HIPCHK(cp(o->rank_sum, c->rank.p, 8ull * C));
    HIPCHK(cp(o->lk10, c->lk.p, 80ull * C));
    HIPCHK(cp(o->depth, c->sites.depth.p, 4ull * C));
    HIPCHK(cp(o->dep_uni, c->sites.nrec.p, 4ull * C));
    HIPCHK(hipStreamSynchronize(c->stream));
    return GSNP_OK;
}

int gsnp_window_recycle(gsnp_ctx* c)
{
    if (!c || !c->win_cap) return fail(GSNP_E_ARG, "window not initialised");
    if (!c->built) {
        int rc = gsnp_window_update(c, 0, 0, nullptr);
        if (rc) return rc;
    }
    c->timer.begin(kStRecycle, c->stream);
    launch_recycle(c->stream, c->site_view(0), GSNP_WIN, (uint32_t)c->p.read_length);
    c->timer.end(kStRecycle, c->stream);
    HIPCHK(hipGetLastError());
    c->ws += GSNP_WIN;
    c->built = false;
    return GSNP_OK;
}

int gsnp_batched_call(gsnp_ctx* c, uint32_t start, uint32_t n, int emit_text, uint64_t* nbytes)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    if ((uint64_t)start + n > c->chr_len) return fail(GSNP_E_ARG, "span outside the chromosome");
    if (start % GSNP_WIN) return fail(GSNP_E_ARG, "span must start on a window boundary");
    int rc = c->span_prepare(n);
    if (rc) return rc;
    const FusedTune& t = c->tune;
    std::vector<uint64_t> key = {c->buf_gen, c->chrom_id, start, n, (uint64_t)c->p.rank_sum_mode,
                                 (uint64_t)c->keep_lk, t.min_class};
    for (int i = 0; i < 6; i++) key.push_back(((uint64_t)t.wpb[i] << 32) | t.max_blocks[i]);
    rc = c->launch_captured(c->graph_slot(key), key, [&] { c->span_enqueue(start, n); });
    if (rc) return rc;
    if (c->p.rank_sum_mode) {
        c->fix_offset = 0;
        rc = c->host_fixups();
        if (rc) return rc;
    }
    rc = c->check_err();
    if (rc) return rc;
    c->ws = start;
    c->built = false;
    if (emit_text) return c->format_text(0, start, n, start + 1, nbytes, true);
    if (nbytes) *nbytes = 0;
    return GSNP_OK;
}

int gsnp_autotune(gsnp_ctx* c, uint32_t start, uint32_t n, int reps, uint32_t* config)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    if (reps < 1) reps = 1;
    int rc = c->span_build(start, n);
    if (rc) return rc;
    SiteArrays a = c->site_view(0);
    const FusedArgs f = c->fused_args(start);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto time_ms = [&](auto&& launch) -> float {
        float best = 3.4e38f;
        for (int r = 0; r < reps; r++) {
            (void)hipMemsetAsync(c->job_count.p, 0, 4, c->stream);
            (void)hipEventRecord(e0, c->stream);
            launch();
            (void)hipEventRecord(e1, c->stream);
            float t = 0.f;
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess) return 3.4e38f;
            best = t < best ? t : best;
        }
        return best;
    };
    static const uint32_t kWpb[4] = {1, 2, 4, 8};
    static const uint32_t kCaps[5] = {512, 1024, 2048, 4096, 8192};
    FusedTune best_tune;
    float best_total = 3.4e38f;
    for (uint32_t mc = 0; mc <= 2; mc++) {  // smallest network: 8, 16 or 32 lanes
        launch_bucket(c->stream, a.nrec, n, c->lists.p, c->win_cap, c->counts.p, 1, nullptr, (int)mc);
        FusedTune t;
        t.min_class = mc;
        float total = 0.f;
        for (uint32_t cls = mc; cls < 6; cls++) {
            float bt = 3.4e38f;
            for (uint32_t w : kWpb)
                for (uint32_t cap : kCaps) {
                    const float ms = time_ms(
                        [&] { launch_fused_one(c->stream, f, c->lists.p, c->win_cap, c->counts.p, n, cls, w, cap); });
                    if (ms < bt) {
                        bt = ms;
                        t.wpb[cls] = w;
                        t.max_blocks[cls] = cap;
                    }
                }
            total += bt;
        }
        if (total < best_total) {
            best_total = total;
            best_tune = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    rc = c->check_err();
    if (rc) return rc;
    c->tune = best_tune;
    if (config) {
        config[0] = best_tune.min_class;
        for (int i = 0; i < 6; i++) config[1 + i] = best_tune.wpb[i], config[7 + i] = best_tune.max_blocks[i];
    }
    return GSNP_OK;
}

int gsnp_sort_likelihood(gsnp_ctx* c, uint32_t* records, uint32_t pitch, const uint32_t* counts, uint32_t n_sites,
                         double* lk10)
{
    if (!c || !records || !counts || !lk10) return fail(GSNP_E_ARG, "null argument");
    if (pitch > kSlots) return fail(GSNP_E_ARG, "pitch above 256");
    int rc = c->alloc_window(n_sites + 1);
    if (rc) return rc;
    std::vector<uint32_t> rows((size_t)n_sites * kSlots, 0xFFFFFFFFu);
    std::vector<int32_t> n(n_sites);
    for (uint32_t i = 0; i < n_sites; i++) {
        if (counts[i] > pitch) return fail(GSNP_E_ARG, "count above pitch");
        n[i] = (int32_t)counts[i];
        std::memcpy(&rows[(size_t)i * kSlots], records + (size_t)i * pitch, 4ull * counts[i]);
    }
    HIPCHK(hipMemcpyAsync(c->sites.rec.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->sites.nrec.p, n.data(), 4ull * n_sites, hipMemcpyHostToDevice, c->stream));
    SiteArrays a = c->site_view(0);
    launch_bucket(c->stream, a.nrec, n_sites, c->lists.p, c->win_cap, c->counts.p, 0, nullptr);
    launch_sort(c->stream, a.rec, a.nrec, c->lists.p, c->win_cap, c->counts.p, n_sites);
    launch_likelihood(c->stream, c->tables(), a, n_sites, c->lk.p, c->err.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(rows.data(), c->sites.rec.p, rows.size() * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(lk10, c->lk.p, 80ull * n_sites, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (uint32_t i = 0; i < n_sites; i++)
        std::memcpy(records + (size_t)i * pitch, &rows[(size_t)i * kSlots], 4ull * counts[i]);
    c->built = false;
    return c->check_err();
}

int gsnp_profile(gsnp_ctx* c, int enable, double* stage_ms, uint64_t* stage_count, int n_stages)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    int rc = c->timer.collect();
    if (rc) return rc;
    for (int i = 0; i < n_stages && i < kNumStages; i++) {
        if (stage_ms) stage_ms[i] = c->timer.ms[i];
        if (stage_count) stage_count[i] = c->timer.count[i];
    }
    for (int i = 0; i < kNumStages; i++) c->timer.ms[i] = 0, c->timer.count[i] = 0;
    c->timer.on = enable != 0;
    return GSNP_OK;
}

int gsnp_stats(gsnp_ctx* c, uint64_t* n_records, int reset)
{
    if (!c) return fail(GSNP_E_ARG, "null context");
    unsigned long long v = 0;
    HIPCHK(hipMemcpyAsync(&v, c->stat.p, 8, hipMemcpyDeviceToHost, c->stream));
    if (reset) HIPCHK(hipMemsetAsync(c->stat.p, 0, 8, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n_records) *n_records = v;
    return GSNP_OK;
}

int gsnp_counters(gsnp_ctx* c, uint64_t* out, int n, int reset)
{
    if (!c || (n > 0 && !out)) return fail(GSNP_E_ARG, "null argument");
    unsigned long long st[2] = {0, 0};  // records called, normal approximations finished on the device
    HIPCHK(hipMemcpyAsync(st, c->stat.p, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const uint64_t v[4] = {st[0], c->fixups_total, c->fixups_normal, st[1]};
    for (int i = 0; i < n && i < 4; i++) out[i] = v[i];
    if (reset) {
        HIPCHK(hipMemsetAsync(c->stat.p, 0, 16, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->fixups_total = c->fixups_normal = 0;
    }
    return GSNP_OK;
}

int gsnp_format_g(double v, char* out)
{
    int n = fmt_g(v, out);
    out[n] = 0;
    return n;
}

}  // extern "C"

namespace gsnp {

// Site matrix of the span [start, start+n) straight from the device read
// store: every read that covers it, no window batches and no carry, with the
// q_sum of window-boundary sites before their own reads kept in head_qs.
int Ctx::span_prepare(uint32_t n)
{
    int rc = alloc_window(n + 2);
    if (rc) return rc;
    const uint32_t nheads = n / GSNP_WIN + 2;
    if (head_qs.n < nheads) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(head_qs.alloc(nheads));
        buf_gen++;
    }
    return GSNP_OK;
}

// The device work of one batched span, up to the rank-sum jobs: site matrix,
// depth classes, fused sort + likelihood + posterior, rank-sum.  Launches
// only (no copies, no memsets, no host waits), so it can be captured.
void Ctx::span_enqueue(uint32_t start, uint32_t n)
{
    launch_fill_u32(stream, head_qs.p, (uint32_t)(4 * (n / GSNP_WIN + 2)), 0u);
    timer.begin(kStSiteBuild, stream);
    launch_site_build(stream, read_view(), 0, n_reads, start, n + 1, INT64_MIN, (int64_t)start + n, 0, site_view(0),
                      head_qs.p, ranges.p);
    timer.end(kStSiteBuild, stream);
    launch_fill_u32(stream, job_count.p, 1, 0u);
    launch_fill_u32(stream, fix_count.p, 1, 0u);
    SiteArrays a = site_view(0);
    timer.begin(kStBucket, stream);
    launch_bucket(stream, a.nrec, n, lists.p, win_cap, counts.p, 1, stat.p, (int)tune.min_class);
    timer.end(kStBucket, stream);
    timer.begin(kStFused, stream);
    launch_fused(stream, fused_args(start), lists.p, win_cap, counts.p, n, tune);
    timer.end(kStFused, stream);
    if (p.rank_sum_mode) {
        Results res = result_view(0);
        timer.begin(kStRank, stream);
        launch_rank(stream, tables(), a, jobs.p, job_count.p, n, res, fixes.p, fix_count.p, stat.p + 1);
        timer.end(kStRank, stream);
    }
}

// Runs enqueue() through a graph captured for `key`, capturing it on first
// use; eagerly when graphs are off or the stage timers are on (their events
// belong to the eager stream).  A capture that fails turns graphs off.
template <class F>
int Ctx::launch_captured(GraphSlot& slot, const std::vector<uint64_t>& key, F enqueue)
{
    if (!use_graphs || timer.on) {
        enqueue();
        HIPCHK(hipGetLastError());
        return GSNP_OK;
    }
    if (!slot.exec || slot.key != key) {
        slot.reset();
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        enqueue();
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (e == hipSuccess && g) e = hipGraphInstantiate(&slot.exec, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess || !slot.exec) {
            (void)hipGetLastError();
            slot.exec = nullptr;
            use_graphs = false;
            enqueue();
            HIPCHK(hipGetLastError());
            return GSNP_OK;
        }
        slot.key = key;
    }
    HIPCHK(hipGraphLaunch(slot.exec, stream));
    return GSNP_OK;
}

constexpr size_t kMaxGraphs = 256;

// The cached graph for `key`, or an empty slot for it.  Graphs captured
// before the last buffer (re)allocation are dropped: their pointers are gone.
GraphSlot& Ctx::graph_slot(const std::vector<uint64_t>& key)
{
    if (graphs_gen != buf_gen) {
        graphs.clear();
        graphs_gen = buf_gen;
    }
    for (auto& g : graphs)
        if (g->key == key) return *g;
    if (graphs.size() >= kMaxGraphs) graphs.erase(graphs.begin());
    graphs.push_back(std::make_unique<GraphSlot>());
    return *graphs.back();
}

int Ctx::span_build(uint32_t start, uint32_t n)
{
    if ((uint64_t)start + n > chr_len) return fail(GSNP_E_ARG, "span outside the chromosome");
    if (start % GSNP_WIN) return fail(GSNP_E_ARG, "span must start on a window boundary");
    int rc = span_prepare(n);
    if (rc) return rc;
    launch_fill_u32(stream, head_qs.p, (uint32_t)(4 * (n / GSNP_WIN + 2)), 0u);
    timer.begin(kStSiteBuild, stream);
    launch_site_build(stream, read_view(), 0, n_reads, start, n + 1, INT64_MIN, (int64_t)start + n, 0, site_view(0),
                      head_qs.p, ranges.p);
    timer.end(kStSiteBuild, stream);
    HIPCHK(hipGetLastError());
    return GSNP_OK;
}

FusedArgs Ctx::fused_args(uint32_t start)
{
    FusedArgs f;
    f.t = tables();
    f.ref = ref_view();
    f.ws = start;
    f.a = site_view(0);
    f.head_qs = head_qs.p;
    f.out = result_view(0);
    f.jobs = jobs.p;
    f.job_count = job_count.p;
    f.err = err.p;
    f.write_lk = keep_lk ? 1 : 0;
    return f;
}

// Bytes of the rows of n sites at most: name, 16 columns and tabs.
static uint64_t text_bound(uint64_t n_sites, size_t name_len) { return n_sites * (name_len + 160) + 64; }

// Formats the rows of sites [offset, offset + n_sites) into `out` (sized by
// text_bound), or into the context's text buffer when out is null.
int Ctx::format_text(uint32_t offset, uint32_t ws_abs, uint32_t n_sites, uint32_t pos_base, uint64_t* nbytes,
                     bool graph, char* out)
{
    const size_t name_len = chr_name.size();
    const uint64_t need = text_bound(n_sites, name_len);
    if (!out && need > text_cap) {
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(text.alloc(need));
        text_cap = need;
        buf_gen++;
    }
    char* dst = out ? out : text.p;
    if (gcount.n < 1) HIPCHK(gcount.alloc(1));
    RowCtx rc;
    rc.name = d_name.p;
    rc.name_len = (int)name_len;
    rc.pos_base = pos_base;
    rc.gidx = gidx.p + offset;
    rc.gtext = gtext.p;
    SiteArrays a = site_view(offset);
    Results res = result_view(offset);
    const uint32_t nb = (n_sites + kRowBlock - 1) / kRowBlock;
    auto enqueue = [&] {
        timer.begin(kStFormat, stream);
        launch_fill_u32(stream, gcount.p, 1, 0u);
        launch_rowlen(stream, ref_view(), ws_abs, a, res, n_sites, rc, gidx.p + offset, gtext.p, gcount.p, rowlen.p,
                      blocksum.p);
        launch_gtext(stream, a, res, n_sites, gtext.p, gcount.p, rowlen.p, blocksum.p);
        launch_scan_blocks(stream, blocksum.p, nb, total.p);
        launch_format(stream, ref_view(), ws_abs, a, res, n_sites, rc, rowlen.p, blocksum.p, dst);
        timer.end(kStFormat, stream);
    };
    if (graph) {
        const std::vector<uint64_t> key = {buf_gen, chrom_id, 1u, offset, ws_abs, n_sites, pos_base,
                                           (uint64_t)(uintptr_t)dst};
        int r = launch_captured(graph_slot(key), key, enqueue);
        if (r) return r;
    } else {
        enqueue();
        HIPCHK(hipGetLastError());
    }
    uint64_t t = 0;
    HIPCHK(hipMemcpyAsync(&t, total.p, 8, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (out) {
        if (nbytes) *nbytes = n_sites ? t : 0;
        return GSNP_OK;
    }
    text_len = n_sites ? t : 0;
    if (nbytes) *nbytes = text_len;
    return GSNP_OK;
}

// Hands the oldest queued span's rows to consume(), once its copy has landed.
static int hand_over_oldest(TextPipe& tp, const Ctx::TextSink& consume)
{
    const int o = (tp.cur - tp.pending + TextPipe::kBufs) % TextPipe::kBufs;
    HIPCHK(hipEventSynchronize(tp.ev_done[o]));
    tp.pending--;
    return consume(tp.host[o], tp.len[o]);
}

int Ctx::text_to_host(uint32_t start, uint32_t n, uint64_t* nbytes, const TextSink& consume)
{
    TextPipe& tp = pipe;
    if (!tp.copy) {
        HIPCHK(hipStreamCreateWithFlags(&tp.copy, hipStreamNonBlocking));
        for (hipEvent_t& e : tp.ev_done) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&tp.ev_fmt, hipEventDisableTiming));
    }
    // buffer cur last held the span kBufs spans back: hand it on first
    if (tp.pending == TextPipe::kBufs) {
        int rc = hand_over_oldest(tp, consume);
        if (rc) return rc;
    }
    const int i = tp.cur;
    const uint64_t need = text_bound(n, chr_name.size());
    if (tp.dev[i].n < need) {
        // nothing may still read the old buffer: its copy, or a graph that writes it
        HIPCHK(hipStreamSynchronize(tp.copy));
        HIPCHK(hipStreamSynchronize(stream));
        HIPCHK(tp.dev[i].alloc(need + need / 8));
        buf_gen++;
    }
    // the copy of the span formatted into this buffer kBufs spans ago reads it
    if (tp.recorded[i]) HIPCHK(hipStreamWaitEvent(stream, tp.ev_done[i], 0));
    uint64_t len = 0;
    int rc = format_text(0, start, n, start + 1, &len, true, tp.dev[i].p);
    if (rc) return rc;
    if (tp.host_cap[i] < len) {  // that span has been handed on, and its copy has landed
        if (tp.host[i]) HIPCHK(hipHostFree(tp.host[i]));
        tp.host[i] = nullptr;
        tp.host_cap[i] = 0;
        const uint64_t cap = len + len / 8 + 4096;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&tp.host[i]), cap, hipHostMallocDefault));
        tp.host_cap[i] = cap;
    }
    HIPCHK(hipEventRecord(tp.ev_fmt, stream));
    HIPCHK(hipStreamWaitEvent(tp.copy, tp.ev_fmt, 0));
    if (len) HIPCHK(hipMemcpyAsync(tp.host[i], tp.dev[i].p, len, hipMemcpyDeviceToHost, tp.copy));
    HIPCHK(hipEventRecord(tp.ev_done[i], tp.copy));
    tp.recorded[i] = true;
    tp.len[i] = len;
    tp.cur = (i + 1) % TextPipe::kBufs;
    tp.pending++;
    if (nbytes) *nbytes = len;
    return GSNP_OK;
}

int Ctx::reset_text()
{
    TextPipe& tp = pipe;
    if (tp.copy) HIPCHK(hipStreamSynchronize(tp.copy));
    tp.pending = 0;
    tp.cur = 0;
    return GSNP_OK;
}

int Ctx::window_text_to_host(const TextSink& consume)
{
    TextPipe& tp = pipe;
    if (tp.copy) HIPCHK(hipStreamSynchronize(tp.copy));
    if (tp.host_cap[0] < text_len) {
        if (tp.host[0]) HIPCHK(hipHostFree(tp.host[0]));
        tp.host[0] = nullptr;
        tp.host_cap[0] = 0;
        const uint64_t cap = text_len + text_len / 8 + 4096;
        HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&tp.host[0]), cap, hipHostMallocDefault));
        tp.host_cap[0] = cap;
    }
    if (text_len) HIPCHK(hipMemcpyAsync(tp.host[0], text.p, text_len, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return consume(tp.host[0], text_len);
}

int Ctx::finish_text(const TextSink& consume)
{
    TextPipe& tp = pipe;
    while (tp.pending) {  // the spans still queued, in order
        int rc = hand_over_oldest(tp, consume);
        if (rc) return rc;
    }
    return GSNP_OK;
}

}  // namespace gsnp
