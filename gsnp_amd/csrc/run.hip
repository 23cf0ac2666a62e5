// run.hip -- whole-file pipeline: FASTA + SOAP (+ known SNPs) to .cns text or
// the -C column files, chromosome by chromosome (main + Call_win::new_soap2cns,
// cpp/main.cc:91-429, cpp/call_genotype.cc:422-669).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "cols.h"
#include "cwriter.h"
#include "engine.h"
#include "inputs.h"

namespace gsnp {

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct FileCloser {
    void operator()(FILE* f) const
    {
        if (f) std::fclose(f);
    }
};

struct CtxDeleter {
    void operator()(gsnp_ctx* c) const { gsnp_destroy(c); }
};

}  // namespace

// One calling pass over a chromosome already loaded into the context.
// batched = 0 runs the per-window path (init, update, likelihood, posterior,
// recycle per window, cpp/call_genotype.cc:474-644); batched = k calls spans
// of k windows straight from the device read store.  After each window
// (per-window path) or span (batched path) the results are handed to
// on_results(offset, span_start, span_sites) and, with emit_text, the span's
// rows are formatted in device memory and their size given to on_text.
template <class OnResults, class OnText>
int call_chromosome(gsnp_ctx* ctx, const std::vector<int32_t>& pos, uint32_t N, int batched, bool emit_text,
                    OnResults on_results, OnText on_text)
{
    int rc;
    const uint32_t nwin = N / GSNP_WIN + 1;
    if (batched) {
        const uint32_t per = (uint32_t)batched;
        for (uint32_t w0 = 0; w0 < nwin; w0 += per) {
            const uint32_t ws = w0 * GSNP_WIN;
            const uint32_t we = (uint32_t)std::min<uint64_t>((uint64_t)(w0 + per) * GSNP_WIN, N);
            const uint32_t n = we > ws ? we - ws : 0;
            uint64_t nb = 0;
            rc = gsnp_batched_call(ctx, ws, n, emit_text, &nb);
            if (rc) return rc;
            rc = on_results(0u, ws, n);
            if (rc) return rc;
            // a chromosome whose length is a multiple of the window ends with an
            // empty window (cpp/call_genotype.cc:636-644); report it on its own
            const uint32_t last = N / GSNP_WIN;
            if (N % GSNP_WIN == 0 && n > 0 && last >= w0 && last < w0 + per) {
                rc = on_results(0u, N, 0u);
                if (rc) return rc;
            }
            if (emit_text) {
                rc = on_text(nb);
                if (rc) return rc;
            }
        }
        return GSNP_OK;
    }
    rc = gsnp_window_init(ctx, 0);
    if (rc) return rc;
    uint64_t r = 0;
    for (uint32_t w = 0; w < nwin; w++) {
        const uint32_t ws = w * GSNP_WIN;
        const uint32_t C = (w < N / GSNP_WIN) ? GSNP_WIN : N % GSNP_WIN;
        const uint64_t r1 = (uint64_t)(std::lower_bound(pos.begin(), pos.end(), (int64_t)ws + GSNP_WIN) - pos.begin());
        rc = gsnp_window_update(ctx, r, r1, nullptr);
        if (rc) return rc;
        r = r1;
        rc = gsnp_window_likelihood(ctx, C);
        if (!rc) rc = gsnp_window_posterior(ctx, C);
        if (rc) return rc;
        rc = on_results(0u, ws, C);
        if (rc) return rc;
        if (emit_text) {
            uint64_t nb = 0;
            rc = gsnp_window_text(ctx, C, &nb);
            if (!rc) rc = on_text(nb);
            if (rc) return rc;
        }
        rc = gsnp_window_recycle(ctx);
        if (rc) return rc;
    }
    return GSNP_OK;
}

}  // namespace gsnp

struct gsnp_cns_reader {
    gsnp::ColumnReader rd;
    gsnp::CnsWindow w;
};

struct gsnp_job {
    gsnp_params p{};
    gsnp::Genome g;
    std::vector<gsnp::ReadBlock> blocks;
    std::vector<double> pm;
    gsnp_ctx* ctx = nullptr;
    int cur = -1;
    ~gsnp_job()
    {
        if (ctx) gsnp_destroy(ctx);
    }
};

namespace gsnp {

int job_select(gsnp_job* job, int index);

int job_open(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
             const gsnp_params& p, std::unique_ptr<gsnp_job>& job)
{
    std::string err;
    job.reset(new gsnp_job());
    job->p = p;
    int rc = load_genome(fasta, dbsnp, job->g, err);
    if (rc) return fail(rc, err);
    rc = load_alignments(soap, p, job->g, job->blocks, err);
    if (rc) return fail(rc, err);
    job->pm.assign((size_t)256 * 256 * 16, 1.0);
    if (matrix_in && matrix_in[0]) {
        rc = matrix_read(matrix_in, job->pm.data(), err);
        if (rc) return fail(rc, err);
        rc = create_ctx(&p, job->pm.data(), &job->ctx);
        if (rc) return rc;
    } else {
        // pass 1 on the device: each chromosome's reads are uploaded and
        // counted, the last one stays selected for calling
        rc = create_ctx(&p, nullptr, &job->ctx);
        if (rc) return rc;
        std::vector<uint64_t> cnt(kCalibBins, 0);
        for (int i = 0; i < (int)job->blocks.size(); i++) {
            rc = job_select(job.get(), i);
            if (!rc) rc = gsnp_calib_count(job->ctx, cnt.data());
            if (rc) return rc;
        }
        matrix_from_counts(cnt.data(), p, job->pm.data());
        rc = job->ctx->set_matrix(job->pm.data());
        if (rc) return rc;
    }
    if (matrix_out && matrix_out[0]) {
        rc = matrix_write(matrix_out, job->pm.data(), p, err);
        if (rc) return fail(rc, err);
    }
    return GSNP_OK;
}

int job_select(gsnp_job* job, int index)
{
    if (index < 0 || index >= (int)job->blocks.size()) return fail(GSNP_E_ARG, "chromosome index out of range");
    if (job->cur == index) return GSNP_OK;  // already on the device
    const ReadBlock& b = job->blocks[index];
    const Chrom& chr = job->g.chr[b.chr];
    int rc = gsnp_set_chromosome(job->ctx, chr.name.c_str(), chr.code.data(), (uint32_t)chr.code.size(),
                                 chr.snp_pos.data(), chr.snp.data(), (uint32_t)chr.snp.size());
    if (rc) return rc;
    gsnp_read_batch rb;
    rb.n = b.size();
    rb.pitch = b.pitch;
    rb.pos = b.pos.data();
    rb.len_hit = b.len_hit.data();
    rb.strand = b.strand.data();
    rb.seq = b.seq.data();
    rb.qual = b.qual.data();
    rc = gsnp_load_reads(job->ctx, &rb);
    if (rc) return rc;
    job->cur = index;
    return GSNP_OK;
}

int run_pipeline(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
                 const gsnp_params& p, const char* out, int compress, int batched)
{
    std::unique_ptr<gsnp_job> job;
    int rc = job_open(fasta, soap, dbsnp, matrix_in, matrix_out, p, job);
    if (rc) return rc;
    std::string err;
    std::unique_ptr<FILE, FileCloser> fo(std::fopen(out, "w"));
    if (!fo) return fail(GSNP_E_IO, std::string("cannot create ") + out);
    CWriter cw;
    if (compress && cw.open(out, err)) return fail(GSNP_E_IO, err);
    std::vector<char> host_text;
    gsnp_ctx* ctx = job->ctx;
    for (int bi = 0; bi < (int)job->blocks.size(); bi++) {
        rc = job_select(job.get(), bi);
        if (rc) return rc;
        const ReadBlock& b = job->blocks[bi];
        const Chrom& chr = job->g.chr[b.chr];
        const uint32_t N = (uint32_t)chr.code.size();
        auto on_results = [&](uint32_t offset, uint32_t ws, uint32_t n) -> int {
            if (!compress) return GSNP_OK;
            if (n == 0) {  // the empty window after a chromosome that fills its last window
                cw.write_empty_window(chr.name, ws);
                return GSNP_OK;
            }
            // one -C chunk per reference window of the span, encoded on the device
            return encode_columns(*ctx, offset, ws, n, [&](const ColumnPass& pass) -> int {
                cw.write_chunks(chr.name, pass);
                return GSNP_OK;
            });
        };
        auto on_text = [&](uint64_t nb) -> int {
            host_text.resize(nb);
            int r2 = gsnp_fetch_text(ctx, host_text.data(), nb, &nb);
            if (r2) return r2;
            std::fwrite(host_text.data(), 1, nb, fo.get());
            return GSNP_OK;
        };
        rc = call_chromosome(ctx, b.pos, N, batched, !compress, on_results, on_text);
        if (rc) return rc;
    }
    if (compress && cw.close(err)) return fail(GSNP_E_IO, err);
    return GSNP_OK;
}

}  // namespace gsnp

extern "C" {

int gsnp_run(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
             const gsnp_params* p, const char* out, int compress, int batched)
{
    if (!fasta || !soap || !p || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::run_pipeline(fasta, soap, dbsnp, matrix_in, matrix_out, *p, out, compress, batched);
}

int gsnp_job_open(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const gsnp_params* p,
                  gsnp_job** out)
{
    if (!fasta || !soap || !p || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    std::unique_ptr<gsnp_job> job;
    int rc = gsnp::job_open(fasta, soap, dbsnp, matrix_in, nullptr, *p, job);
    if (rc) return rc;
    *out = job.release();
    return GSNP_OK;
}

int gsnp_job_count(gsnp_job* job) { return job ? (int)job->blocks.size() : 0; }

int gsnp_job_info(gsnp_job* job, int index, char* name, uint64_t cap, uint32_t* length, uint64_t* n_reads)
{
    if (!job || index < 0 || index >= (int)job->blocks.size()) return gsnp::fail(GSNP_E_ARG, "bad chromosome index");
    const gsnp::ReadBlock& b = job->blocks[index];
    const gsnp::Chrom& c = job->g.chr[b.chr];
    if (name && cap) {
        std::snprintf(name, cap, "%s", c.name.c_str());
    }
    if (length) *length = (uint32_t)c.code.size();
    if (n_reads) *n_reads = b.size();
    return GSNP_OK;
}

int gsnp_job_select(gsnp_job* job, int index)
{
    if (!job) return gsnp::fail(GSNP_E_ARG, "null job");
    return gsnp::job_select(job, index);
}

int gsnp_job_call(gsnp_job* job, int batched, int emit_text, uint64_t* nbytes)
{
    if (!job || job->cur < 0) return gsnp::fail(GSNP_E_ARG, "no chromosome selected");
    const gsnp::ReadBlock& b = job->blocks[job->cur];
    const uint32_t N = (uint32_t)job->g.chr[b.chr].code.size();
    uint64_t total = 0;
    auto on_results = [](uint32_t, uint32_t, uint32_t) -> int { return GSNP_OK; };
    auto on_text = [&](uint64_t nb) -> int {
        total += nb;
        return GSNP_OK;
    };
    int rc = gsnp::call_chromosome(job->ctx, b.pos, N, batched, emit_text != 0, on_results, on_text);
    if (nbytes) *nbytes = total;
    return rc;
}

int gsnp_job_columns(gsnp_job* job, int batched, uint64_t* nbytes)
{
    if (!job || job->cur < 0) return gsnp::fail(GSNP_E_ARG, "no chromosome selected");
    const gsnp::ReadBlock& b = job->blocks[job->cur];
    const uint32_t N = (uint32_t)job->g.chr[b.chr].code.size();
    uint64_t total = 0;
    auto on_results = [&](uint32_t offset, uint32_t ws, uint32_t n) -> int {
        if (n == 0) return GSNP_OK;
        return gsnp::encode_columns(*job->ctx, offset, ws, n, [&](const gsnp::ColumnPass& pass) -> int {
            for (int f = 0; f < 5; f++) total += pass.bytes[f];
            return GSNP_OK;
        });
    };
    auto on_text = [](uint64_t) -> int { return GSNP_OK; };
    int rc = gsnp::call_chromosome(job->ctx, b.pos, N, batched, false, on_results, on_text);
    if (nbytes) *nbytes = total;
    return rc;
}

gsnp_ctx* gsnp_job_context(gsnp_job* job) { return job ? job->ctx : nullptr; }

void gsnp_job_close(gsnp_job* job) { delete job; }

int gsnp_decompress(const char* prefix, const char* out)
{
    std::string err;
    int rc = gsnp::decompress_files(prefix, out, err);
    return rc ? gsnp::fail(rc == 1 ? GSNP_E_IO : GSNP_E_FORMAT, err) : GSNP_OK;
}

int gsnp_cns_open(const char* prefix, gsnp_cns_reader** out)
{
    if (!prefix || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    std::unique_ptr<gsnp_cns_reader> r(new gsnp_cns_reader());
    std::string err;
    int rc = r->rd.open(prefix, err);
    if (rc) return gsnp::fail(rc == 1 ? GSNP_E_IO : GSNP_E_FORMAT, err);
    *out = r.release();
    return GSNP_OK;
}

int gsnp_cns_next(gsnp_cns_reader* r, gsnp_cns_window* w, int* has_window)
{
    if (!r || !w || !has_window) return gsnp::fail(GSNP_E_ARG, "null argument");
    std::string err;
    bool got = false;
    int rc = r->rd.next(r->w, got, err);
    *has_window = got ? 1 : 0;
    if (rc) return gsnp::fail(GSNP_E_FORMAT, err);
    if (!got) return GSNP_OK;
    const gsnp::CnsWindow& c = r->w;
    w->chr = c.chr.c_str();
    w->start = c.start;
    w->n = c.n;
    w->ref = c.c3.data(), w->cns = c.c4.data(), w->q_cns = c.c5.data(), w->base1 = c.c6.data();
    w->avg_q1 = c.c7.data(), w->count_uni1 = c.c8.data(), w->count_all1 = c.c9.data(), w->base2 = c.c10.data();
    w->avg_q2 = c.c11.data(), w->count_uni2 = c.c12.data(), w->count_all2 = c.c13.data(), w->depth = c.c14.data();
    w->rank_sum = c.c15.data(), w->copy_num = c.c16.data(), w->dbsnp = c.c17.data();
    return GSNP_OK;
}

void gsnp_cns_close(gsnp_cns_reader* r) { delete r; }

int gsnp_matrix_gen(const char* fasta, const char* soap, const char* dbsnp, const gsnp_params* p, double* pm)
{
    if (!fasta || !soap || !p || !pm) return gsnp::fail(GSNP_E_ARG, "null argument");
    std::unique_ptr<gsnp_job> job;
    int rc = gsnp::job_open(fasta, soap, dbsnp, nullptr, nullptr, *p, job);
    if (rc) return rc;
    std::copy(job->pm.begin(), job->pm.end(), pm);
    return GSNP_OK;
}

}  // extern "C"
