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
#include "synth.h"

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

// Text modes of a calling pass: no rows, rows formatted in device memory, or
// rows copied on to host memory.
enum TextMode : int { kTextNone = 0, kTextDevice = 1, kTextHost = 2 };

// One calling pass over a chromosome already loaded into the context.
// batched = 0 runs the per-window path (init, update, likelihood, posterior,
// recycle per window, cpp/call_genotype.cc:474-644); batched = k calls spans
// of k windows straight from the device read store.  After each window
// (per-window path) or span (batched path) the results are handed to
// on_results(offset, span_start, span_sites).  With kTextDevice the rows are
// formatted in device memory; with kTextHost they also go to host memory and
// to sink(ptr, len) in order: per window with one synchronous copy (the
// reference's structure), per span through the context's text pipe, whose
// copy of one span runs beside the next span's kernels (the caller ends the
// pass with finish_text).  *text_bytes adds up the rows.
template <class OnResults>
int call_chromosome(gsnp_ctx* ctx, const std::vector<int32_t>& pos, uint32_t N, int batched, int text,
                    OnResults on_results, uint64_t* text_bytes, const Ctx::TextSink& sink,
                    const std::vector<uint64_t>* mask = nullptr)
{
    int rc;
    const uint32_t nwin = N / GSNP_WIN + 1;
    // with -T regions only the windows they touch are called, each span a run
    // of consecutive called windows (the per-window path then calls one-window
    // spans: a window after a skipped one has no carried tail)
    if (batched || mask) {
        const uint64_t per = batched > 0 ? (uint64_t)batched : 1u;
        for (uint32_t w0 = 0; w0 < nwin;) {
            if (mask && !window_called(*mask, w0)) {
                w0++;
                continue;
            }
            uint32_t w1 = w0 + 1;
            while (w1 < nwin && w1 - w0 < per && (!mask || window_called(*mask, w1))) w1++;
            const uint32_t ws = w0 * GSNP_WIN;
            const uint32_t we = (uint32_t)std::min<uint64_t>((uint64_t)w1 * GSNP_WIN, N);
            const uint32_t n = we > ws ? we - ws : 0;
            uint64_t nb = 0;
            rc = gsnp_batched_call(ctx, ws, n, text == kTextDevice, &nb);
            if (rc) return rc;
            rc = on_results(0u, ws, n);
            if (rc) return rc;
            // a chromosome whose length is a multiple of the window ends with an
            // empty window (cpp/call_genotype.cc:636-644); report it on its own
            const uint32_t last = N / GSNP_WIN;
            if (N % GSNP_WIN == 0 && n > 0 && last >= w0 && last < w1) {
                rc = on_results(0u, N, 0u);
                if (rc) return rc;
            }
            if (text == kTextHost) {
                rc = ctx->text_to_host(ws, n, &nb, sink);
                if (rc) return rc;
            }
            if (text_bytes) *text_bytes += nb;
            w0 = w1;
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
        if (text != kTextNone) {
            uint64_t nb = 0;
            rc = gsnp_window_text(ctx, C, &nb);
            if (!rc && text == kTextHost) rc = ctx->window_text_to_host(sink);
            if (rc) return rc;
            if (text_bytes) *text_bytes += nb;
        }
        rc = gsnp_window_recycle(ctx);
        if (rc) return rc;
    }
    return GSNP_OK;
}

// The five chunks of a window with no sites (the window after a chromosome
// that fills its last window, cpp/call_genotype.cc:636-644): an all-N
// .3_4_6 header and the N indicator of the four other files.
struct EmptyChunks {
    uint64_t h346[7] = {1, 0, 0, 0, 0, 0, 0};
    int32_t nind = 1;
    const uint8_t* chunk[5];
    uint64_t bytes[5] = {sizeof(h346), 4, 4, 4, 4};
    EmptyChunks()
    {
        chunk[0] = reinterpret_cast<const uint8_t*>(h346);
        for (int f = 1; f < 5; f++) chunk[f] = reinterpret_cast<const uint8_t*>(&nind);
    }
};

// Window w's chunk in each file of an encoder pass.
inline void pass_window(const ColumnPass& pass, uint32_t w, const uint8_t* chunk[5], uint64_t bytes[5])
{
    for (int f = 0; f < 5; f++) {
        const uint64_t a = pass.offsets[(size_t)w * 5 + f];
        const uint64_t b = w + 1 < pass.n_windows ? pass.offsets[(size_t)(w + 1) * 5 + f] : pass.bytes[f];
        chunk[f] = pass.file[f] + a;
        bytes[f] = b - a;
    }
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
    // every chromosome's device inputs stay resident once uploaded: the ones
    // not selected are parked here (the selected one's slot is empty)
    std::vector<gsnp::ChromState> parked;
    // -T: per chromosome of g, the windows to call (gsnp_job_set_regions)
    bool use_regions = false;
    std::vector<std::vector<uint64_t>> win_mask;
    const std::vector<uint64_t>* mask(int chr) const { return use_regions ? &win_mask[chr] : nullptr; }
    ~gsnp_job()
    {
        parked.clear();
        if (ctx) gsnp_destroy(ctx);
    }
};

namespace gsnp {

int job_select(gsnp_job* job, int index);

// The calibration matrix (read, or counted on the device over every
// chromosome) and the context of a job whose genome and reads are loaded.
int job_prepare(gsnp_job* job, const char* matrix_in, const char* matrix_out);

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
    return job_prepare(job.get(), matrix_in, matrix_out);
}

int job_prepare(gsnp_job* job, const char* matrix_in, const char* matrix_out)
{
    std::string err;
    const gsnp_params& p = job->p;
    int rc;
    job->parked.clear();
    job->parked.resize(job->blocks.size());
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
            rc = job_select(job, i);
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
    if (job->cur >= 0) job->ctx->swap_chrom(job->parked[job->cur]);  // park the selected chromosome
    job->cur = -1;  // nothing is selected until the chromosome is swapped in or uploaded
    if (job->parked[index].chrom_id) {
        job->ctx->swap_chrom(job->parked[index]);
        job->cur = index;
        return GSNP_OK;
    }
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

// Every chromosome of an open job called and written to `out` (.cns text) or,
// with compress, to the -C files out.1_2 ... out.15_17.
int job_write(gsnp_job* job, const char* out, int compress, int batched)
{
    int rc;
    std::string err;
    std::unique_ptr<FILE, FileCloser> fo(std::fopen(out, "w"));
    if (!fo) return fail(GSNP_E_IO, std::string("cannot create ") + out);
    CWriter cw;
    if (compress && cw.open(out, err)) return fail(GSNP_E_IO, err);
    gsnp_ctx* ctx = job->ctx;
    // rows land in pinned host memory and go to the file in order
    const Ctx::TextSink sink = [&](const char* rows, uint64_t len) -> int {
        if (len && std::fwrite(rows, 1, len, fo.get()) != len) return fail(GSNP_E_IO, std::string("cannot write ") + out);
        return GSNP_OK;
    };
    rc = ctx->reset_text();
    if (rc) return rc;
    for (int bi = 0; bi < (int)job->blocks.size(); bi++) {
        rc = job_select(job, bi);
        if (rc) return rc;
        const ReadBlock& b = job->blocks[bi];
        const Chrom& chr = job->g.chr[b.chr];
        const uint32_t N = (uint32_t)chr.code.size();
        auto on_results = [&](uint32_t offset, uint32_t ws, uint32_t n) -> int {
            if (!compress) return GSNP_OK;
            if (n == 0) {  // the empty window after a chromosome that fills its last window
                const int r = finish_columns(*ctx);  // the queued chunks go first
                if (r) return r;
                return cw.write_empty_window(chr.name, ws, err) ? fail(GSNP_E_IO, err) : GSNP_OK;
            }
            // one -C chunk per reference window of the span, encoded on the device
            return encode_columns(*ctx, offset, ws, n, [&](const ColumnPass& pass) -> int {
                return cw.write_chunks(chr.name, pass, err) ? fail(GSNP_E_IO, err) : GSNP_OK;
            });
        };
        rc = call_chromosome(ctx, b.pos, N, batched, compress ? kTextNone : kTextHost, on_results, nullptr, sink,
                             job->mask(b.chr));
        if (!rc && compress) rc = finish_columns(*ctx);  // the sink refers to this chromosome
        if (rc) {
            drop_columns(*ctx);
            return rc;
        }
    }
    if (!compress) {
        rc = ctx->finish_text(sink);
        if (rc) return rc;
    }
    if (compress && cw.close(err)) return fail(GSNP_E_IO, err);
    if (std::fclose(fo.release()) != 0) return fail(GSNP_E_IO, std::string("cannot write ") + out);
    return GSNP_OK;
}

int run_pipeline(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
                 const gsnp_params& p, const char* out, int compress, int batched)
{
    std::unique_ptr<gsnp_job> job;
    int rc = job_open(fasta, soap, dbsnp, matrix_in, matrix_out, p, job);
    if (rc) return rc;
    return job_write(job.get(), out, compress, batched);
}

}  // namespace gsnp

extern "C" {

int gsnp_run(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
             const gsnp_params* p, const char* out, int compress, int batched)
{
    if (!fasta || !soap || !p || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded(
        [&] { return gsnp::run_pipeline(fasta, soap, dbsnp, matrix_in, matrix_out, *p, out, compress, batched); });
}

int gsnp_job_open(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const gsnp_params* p,
                  gsnp_job** out)
{
    if (!fasta || !soap || !p || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded([&] {
        std::unique_ptr<gsnp_job> job;
        int rc = gsnp::job_open(fasta, soap, dbsnp, matrix_in, nullptr, *p, job);
        if (rc) return rc;
        *out = job.release();
        return GSNP_OK;
    });
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
    if (emit_text < 0 || emit_text > 2) return gsnp::fail(GSNP_E_ARG, "emit_text must be 0, 1 or 2");
    uint64_t total = 0;
    auto on_results = [](uint32_t, uint32_t, uint32_t) -> int { return GSNP_OK; };
    const gsnp::Ctx::TextSink landed = [](const char*, uint64_t) -> int { return GSNP_OK; };
    int rc = job->ctx->reset_text();
    if (!rc)
        rc = gsnp::call_chromosome(job->ctx, b.pos, N, batched, emit_text, on_results, &total, landed, job->mask(b.chr));
    if (!rc && emit_text == gsnp::kTextHost) rc = job->ctx->finish_text(landed);
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
    int rc = gsnp::call_chromosome(job->ctx, b.pos, N, batched, gsnp::kTextNone, on_results, nullptr, nullptr,
                                   job->mask(b.chr));
    if (!rc) rc = gsnp::finish_columns(*job->ctx);
    if (rc) gsnp::drop_columns(*job->ctx);
    if (nbytes) *nbytes = total;
    return rc;
}

int gsnp_job_columns_to(gsnp_job* job, int batched, gsnp_chunk_fn fn, void* user, uint64_t* nbytes)
{
    if (!job || job->cur < 0) return gsnp::fail(GSNP_E_ARG, "no chromosome selected");
    if (!fn) return gsnp::fail(GSNP_E_ARG, "null callback");
    return gsnp::guarded([&] {
        const gsnp::ReadBlock& b = job->blocks[job->cur];
        const std::string& name = job->g.chr[b.chr].name;
        const uint32_t N = (uint32_t)job->g.chr[b.chr].code.size();
        uint64_t total = 0;
        auto hand = [&](uint64_t start, uint32_t n, const uint8_t* const* chunk, const uint64_t* bytes) -> int {
            for (int f = 0; f < 5; f++) total += bytes[f];
            const int r = fn(user, name.c_str(), start, n, chunk, bytes);
            return r ? gsnp::fail(GSNP_E_IO, "the chunk callback returned " + std::to_string(r)) : GSNP_OK;
        };
        auto on_results = [&](uint32_t offset, uint32_t ws, uint32_t n) -> int {
            if (n == 0) {
                const int r = gsnp::finish_columns(*job->ctx);  // the queued chunks go first
                if (r) return r;
                const gsnp::EmptyChunks e;
                return hand((uint64_t)ws + 1, 0, e.chunk, e.bytes);
            }
            return gsnp::encode_columns(*job->ctx, offset, ws, n, [&](const gsnp::ColumnPass& pass) -> int {
                for (uint32_t w = 0; w < pass.n_windows; w++) {
                    const uint8_t* chunk[5];
                    uint64_t bytes[5];
                    gsnp::pass_window(pass, w, chunk, bytes);
                    const int r = hand(pass.starts[w], pass.sizes[w], chunk, bytes);
                    if (r) return r;
                }
                return GSNP_OK;
            });
        };
        int rc = gsnp::call_chromosome(job->ctx, b.pos, N, batched, gsnp::kTextNone, on_results, nullptr, nullptr,
                                       job->mask(b.chr));
        if (!rc) rc = gsnp::finish_columns(*job->ctx);
        if (rc) gsnp::drop_columns(*job->ctx);
        if (nbytes) *nbytes = total;
        return rc;
    });
}

int gsnp_window_columns(gsnp_ctx* c, uint32_t call_length, uint8_t* const* dst, const uint64_t* cap, uint64_t* sizes)
{
    if (!c || !sizes) return gsnp::fail(GSNP_E_ARG, "null argument");
    if (!c->win_cap) return gsnp::fail(GSNP_E_ARG, "window not initialised");
    if (call_length > GSNP_WIN) return gsnp::fail(GSNP_E_ARG, "call length above the window size");
    if (dst && !cap) return gsnp::fail(GSNP_E_ARG, "destinations without capacities");
    return gsnp::guarded([&] {
        auto deliver = [&](const uint8_t* const* chunk, const uint64_t* bytes) -> int {
            for (int f = 0; f < 5; f++) sizes[f] = bytes[f];
            if (!dst) return GSNP_OK;
            for (int f = 0; f < 5; f++)
                if (bytes[f] && (!dst[f] || cap[f] < bytes[f]))
                    return gsnp::fail(GSNP_E_ARG, "chunk of file " + std::to_string(f) + " needs " +
                                                      std::to_string(bytes[f]) + " bytes");
            for (int f = 0; f < 5; f++)
                if (bytes[f]) std::memcpy(dst[f], chunk[f], bytes[f]);
            return GSNP_OK;
        };
        if (call_length == 0) {
            const gsnp::EmptyChunks e;
            return deliver(e.chunk, e.bytes);
        }
        int rc = gsnp::encode_columns(*c, 0, c->ws, call_length, [&](const gsnp::ColumnPass& pass) -> int {
            const uint8_t* chunk[5];
            uint64_t bytes[5];
            gsnp::pass_window(pass, 0, chunk, bytes);
            return deliver(chunk, bytes);
        });
        if (!rc) rc = gsnp::finish_columns(*c);  // the window's chunks are the caller's on return
        if (rc) gsnp::drop_columns(*c);
        return rc;
    });
}

int gsnp_job_synth(int config, unsigned long long seed, double scale, const gsnp_params* p, const char* matrix_in,
                   gsnp_job** out, int* read_len)
{
    if (!p || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded([&] {
        gsnp::SynthSet s;
        gsnp::synth_generate(config, seed, scale, nullptr, gsnp::synth_threads(), s);
        std::unique_ptr<gsnp_job> job(new gsnp_job());
        job->p = *p;
        job->p.read_length = s.read_len;
        const uint32_t L = (uint32_t)s.read_len;
        for (gsnp::SynthContig& ct : s.contigs) {
            for (const gsnp::SynthRead& r : ct.reads)
                if (r.indel || r.hit == 0)
                    return gsnp::fail(GSNP_E_ARG, "in-memory synthetic jobs take configurations without indel edits");
            const int ci = (int)job->g.chr.size();
            job->g.index[ct.name] = ci;
            job->g.chr.emplace_back();
            gsnp::Chrom& c = job->g.chr.back();
            c.name = ct.name;
            c.code.swap(ct.ref);  // A0 C1 T2 G3 and 7 for N: the FASTA binarization (ch >> 1) & 7
            for (const gsnp::SynthSnp& t : ct.snps) {
                gsnp_snp g{};
                g.hapmap = t.hapmap;
                g.validated = t.validated;
                g.indel = t.indel;
                std::memcpy(g.freq, t.freq, sizeof g.freq);
                c.snp_pos.push_back(t.pos);
                c.snp.push_back(g);
                c.code[t.pos] |= 8;
            }
            if (ct.reads.empty()) continue;
            job->blocks.emplace_back();
            gsnp::ReadBlock& b = job->blocks.back();
            b.chr = ci;
            b.pitch = L;
            const size_t n = ct.reads.size();
            b.pos.resize(n);
            b.len_hit.resize(n);
            b.strand.resize(n);
            for (size_t i = 0; i < n; i++) {
                b.pos[i] = ct.reads[i].pos;
                b.len_hit[i] = (int32_t)((L << 16) | ct.reads[i].hit);
                b.strand[i] = ct.reads[i].strand ? '-' : '+';
            }
            std::vector<gsnp::SynthRead>().swap(ct.reads);
            b.seq.swap(ct.seq);
            b.qual.swap(ct.qual);
        }
        int rc = gsnp::job_prepare(job.get(), matrix_in, nullptr);
        if (rc) return rc;
        if (read_len) *read_len = s.read_len;
        *out = job.release();
        return GSNP_OK;
    });
}

int gsnp_job_pass(gsnp_job* job, int batched, int emit_text, uint64_t* nbytes, uint64_t* n_sites)
{
    if (!job || !job->ctx) return gsnp::fail(GSNP_E_ARG, "null job");
    if (emit_text < 0 || emit_text > 3) return gsnp::fail(GSNP_E_ARG, "emit_text must be 0, 1, 2 or 3");
    uint64_t total = 0, sites = 0;
    const bool columns = emit_text == 3;
    auto on_results = [&](uint32_t offset, uint32_t ws, uint32_t n) -> int {
        if (!columns || n == 0) return GSNP_OK;
        return gsnp::encode_columns(*job->ctx, offset, ws, n, [&](const gsnp::ColumnPass& pass) -> int {
            for (int f = 0; f < 5; f++) total += pass.bytes[f];
            return GSNP_OK;
        });
    };
    const gsnp::Ctx::TextSink landed = [](const char*, uint64_t) -> int { return GSNP_OK; };
    int rc = job->ctx->reset_text();
    for (int i = 0; !rc && i < (int)job->blocks.size(); i++) {
        rc = gsnp::job_select(job, i);
        if (rc) break;
        const gsnp::ReadBlock& b = job->blocks[i];
        const uint32_t N = (uint32_t)job->g.chr[b.chr].code.size();
        uint32_t per = (uint32_t)batched;
        if (batched < 0) per = N / GSNP_WIN + 1;  // the whole chromosome in one span
        rc = gsnp::call_chromosome(job->ctx, b.pos, N, (int)per, columns ? (int)gsnp::kTextNone : emit_text,
                                   on_results, &total, landed, job->mask(b.chr));
        sites += N;
    }
    if (columns) {  // the last pass's chunks, queued beside the next chromosome's calls until now
        if (!rc) rc = gsnp::finish_columns(*job->ctx);
        if (rc) gsnp::drop_columns(*job->ctx);
    }
    if (!rc && emit_text == gsnp::kTextHost) rc = job->ctx->finish_text(landed);
    if (nbytes) *nbytes = total;
    if (n_sites) *n_sites = sites;
    return rc;
}

int gsnp_job_set_regions(gsnp_job* job, const char* path)
{
    if (!job) return gsnp::fail(GSNP_E_ARG, "null job");
    return gsnp::guarded([&] {
        if (!path || !path[0]) {
            job->use_regions = false;
            job->win_mask.clear();
            return GSNP_OK;
        }
        std::string err;
        std::vector<std::vector<uint64_t>> m;
        int rc = gsnp::load_regions(path, job->g, job->p.read_length, m, err);
        if (rc) return gsnp::fail(rc, err);
        job->win_mask.swap(m);
        job->use_regions = true;
        return GSNP_OK;
    });
}

int gsnp_job_write_matrix(gsnp_job* job, const char* path)
{
    if (!job || !path) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded([&] {
        std::string err;
        int rc = gsnp::matrix_write(path, job->pm.data(), job->p, err);
        return rc ? gsnp::fail(rc, err) : GSNP_OK;
    });
}

int gsnp_job_write(gsnp_job* job, const char* out, int compress, int batched)
{
    if (!job || !job->ctx || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded([&] { return gsnp::job_write(job, out, compress, batched); });
}

gsnp_ctx* gsnp_job_context(gsnp_job* job) { return job ? job->ctx : nullptr; }

void gsnp_job_close(gsnp_job* job) { delete job; }

int gsnp_decompress(const char* prefix, const char* out)
{
    if (!prefix || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded([&] {
        std::string err;
        int rc = gsnp::decompress_files(prefix, out, err);
        return rc ? gsnp::fail(rc == 1 ? GSNP_E_IO : GSNP_E_FORMAT, err) : GSNP_OK;
    });
}

int gsnp_cns_open(const char* prefix, gsnp_cns_reader** out)
{
    if (!prefix || !out) return gsnp::fail(GSNP_E_ARG, "null argument");
    return gsnp::guarded([&] {
        std::unique_ptr<gsnp_cns_reader> r(new gsnp_cns_reader());
        std::string err;
        int rc = r->rd.open(prefix, err);
        if (rc) return gsnp::fail(rc == 1 ? GSNP_E_IO : GSNP_E_FORMAT, err);
        *out = r.release();
        return GSNP_OK;
    });
}

int gsnp_cns_next(gsnp_cns_reader* r, gsnp_cns_window* w, int* has_window)
{
    if (!r || !w || !has_window) return gsnp::fail(GSNP_E_ARG, "null argument");
    std::string err;
    bool got = false;
    *has_window = 0;
    int rc = gsnp::guarded([&] { return r->rd.next(r->w, got, err); });
    if (rc == GSNP_E_NOMEM) return rc;
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
    return gsnp::guarded([&] {
        std::unique_ptr<gsnp_job> job;
        int rc = gsnp::job_open(fasta, soap, dbsnp, nullptr, nullptr, *p, job);
        if (rc) return rc;
        std::copy(job->pm.begin(), job->pm.end(), pm);
        return GSNP_OK;
    });
}

}  // extern "C"
