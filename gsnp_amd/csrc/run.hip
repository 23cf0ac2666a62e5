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
#include "cwriter.h"
#include "engine.h"
#include "inputs.h"

namespace gsnp {

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) return fail(GSNP_E_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device results of sites [offset, offset+C) to host columns.
int fetch_cols(Ctx& c, uint32_t offset, uint32_t ws_abs, uint32_t C, const uint8_t* ref_codes, WindowCols& w)
{
    w.resize(C);
    if (!C) return GSNP_OK;
    std::memcpy(w.ori.data(), ref_codes + ws_abs, C);
    for (uint32_t j = 0; j < C; j++) w.ori[j] &= 0xF;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream);
    };
    HIPCHK(cp(w.base1.data(), c.base1.p + offset, C));
    HIPCHK(cp(w.base2.data(), c.base2.p + offset, C));
    HIPCHK(cp(w.type1.data(), c.type1.p + offset, C));
    HIPCHK(cp(w.q_cns.data(), c.qcns.p + offset, 4ull * C));
    HIPCHK(cp(w.rank.data(), c.rank.p + offset, 8ull * C));
    HIPCHK(cp(w.depth.data(), c.sites.depth.p + offset, 4ull * C));
    HIPCHK(cp(w.repeat.data(), c.sites.repeat.p + offset, 4ull * C));
    HIPCHK(cp(w.qs.data(), c.sites.qs.p + offset, 16ull * C));
    HIPCHK(cp(w.cu.data(), c.sites.cu.p + offset, 16ull * C));
    HIPCHK(cp(w.ca.data(), c.sites.ca.p + offset, 16ull * C));
    HIPCHK(hipStreamSynchronize(c.stream));
    return GSNP_OK;
}

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

int run_pipeline(const char* fasta, const char* soap, const char* dbsnp, const char* matrix_in, const char* matrix_out,
                 const gsnp_params& p, const char* out, int compress, int batched)
{
    std::string err;
    Genome g;
    int rc = load_genome(fasta, dbsnp, g, err);
    if (rc) return fail(rc, err);
    std::vector<ReadBlock> blocks;
    rc = load_alignments(soap, p, g, blocks, err);
    if (rc) return fail(rc, err);
    std::vector<double> pm((size_t)256 * 256 * 16);
    if (matrix_in && matrix_in[0]) {
        rc = matrix_read(matrix_in, pm.data(), err);
        if (rc) return fail(rc, err);
    } else {
        matrix_gen(blocks, g, p, pm.data());
    }
    if (matrix_out && matrix_out[0]) {
        rc = matrix_write(matrix_out, pm.data(), p, err);
        if (rc) return fail(rc, err);
    }
    gsnp_ctx* raw = nullptr;
    rc = gsnp_create(&p, pm.data(), &raw);
    if (rc) return rc;
    std::unique_ptr<gsnp_ctx, CtxDeleter> ctx(raw);
    std::unique_ptr<FILE, FileCloser> fo(std::fopen(out, "w"));
    if (!fo) return fail(GSNP_E_IO, std::string("cannot create ") + out);
    CWriter cw;
    if (compress && cw.open(out, err)) return fail(GSNP_E_IO, err);
    std::vector<char> host_text;
    WindowCols cols;
    for (const ReadBlock& b : blocks) {
        const Chrom& chr = g.chr[b.chr];
        const uint32_t N = (uint32_t)chr.code.size();
        rc = gsnp_set_chromosome(ctx.get(), chr.name.c_str(), chr.code.data(), N, chr.snp_pos.data(), chr.snp.data(),
                                 (uint32_t)chr.snp.size());
        if (rc) return rc;
        gsnp_read_batch rb;
        rb.n = b.size();
        rb.pitch = b.pitch;
        rb.pos = b.pos.data();
        rb.len_hit = b.len_hit.data();
        rb.strand = b.strand.data();
        rb.seq = b.seq.data();
        rb.qual = b.qual.data();
        rc = gsnp_load_reads(ctx.get(), &rb);
        if (rc) return rc;
        const uint32_t nwin = N / GSNP_WIN + 1;
        if (batched) {
            // whole chromosome through the fused path, chunk by chunk of windows
            const uint32_t per = batched > 1 ? (uint32_t)batched : 8u;
            for (uint32_t w0 = 0; w0 < nwin; w0 += per) {
                const uint32_t ws = w0 * GSNP_WIN;
                const uint32_t we = std::min<uint64_t>((uint64_t)(w0 + per) * GSNP_WIN, N);
                const uint32_t n = we > ws ? we - ws : 0;
                uint64_t nb = 0;
                rc = gsnp_batched_call(ctx.get(), ws, n, !compress, &nb);
                if (rc) return rc;
                if (compress) {
                    for (uint32_t w = w0; w < std::min(nwin, w0 + per); w++) {
                        const uint32_t s = w * GSNP_WIN;
                        const uint32_t C = (w < N / GSNP_WIN) ? GSNP_WIN : N % GSNP_WIN;
                        rc = fetch_cols(*ctx, s - ws, s, C, chr.code.data(), cols);
                        if (rc) return rc;
                        cw.write_window(chr.name, s, cols);
                    }
                } else {
                    host_text.resize(nb);
                    rc = gsnp_fetch_text(ctx.get(), host_text.data(), nb, &nb);
                    if (rc) return rc;
                    std::fwrite(host_text.data(), 1, nb, fo.get());
                }
            }
            continue;
        }
        rc = gsnp_window_init(ctx.get(), 0);
        if (rc) return rc;
        uint64_t r = 0;
        for (uint32_t w = 0; w < nwin; w++) {
            const uint32_t ws = w * GSNP_WIN;
            const uint32_t C = (w < N / GSNP_WIN) ? GSNP_WIN : N % GSNP_WIN;
            const uint64_t r1 = (uint64_t)(std::lower_bound(b.pos.begin(), b.pos.end(), (int64_t)ws + GSNP_WIN) - b.pos.begin());
            rc = gsnp_window_update(ctx.get(), r, r1, nullptr);
            if (rc) return rc;
            r = r1;
            rc = gsnp_window_likelihood(ctx.get(), C);
            if (!rc) rc = gsnp_window_posterior(ctx.get(), C);
            if (rc) return rc;
            if (compress) {
                rc = fetch_cols(*ctx, 0, ws, C, chr.code.data(), cols);
                if (rc) return rc;
                cw.write_window(chr.name, ws, cols);
            } else {
                uint64_t nb = 0;
                rc = gsnp_window_text(ctx.get(), C, &nb);
                if (rc) return rc;
                host_text.resize(nb);
                rc = gsnp_fetch_text(ctx.get(), host_text.data(), nb, &nb);
                if (rc) return rc;
                std::fwrite(host_text.data(), 1, nb, fo.get());
            }
            rc = gsnp_window_recycle(ctx.get());
            if (rc) return rc;
        }
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

int gsnp_decompress(const char* prefix, const char* out)
{
    std::string err;
    int rc = gsnp::decompress_files(prefix, out, err);
    return rc ? gsnp::fail(rc == 1 ? GSNP_E_IO : GSNP_E_FORMAT, err) : GSNP_OK;
}

int gsnp_matrix_gen(const char* fasta, const char* soap, const char* dbsnp, const gsnp_params* p, double* pm)
{
    std::string err;
    gsnp::Genome g;
    int rc = gsnp::load_genome(fasta, dbsnp, g, err);
    if (rc) return gsnp::fail(rc, err);
    std::vector<gsnp::ReadBlock> blocks;
    rc = gsnp::load_alignments(soap, *p, g, blocks, err);
    if (rc) return gsnp::fail(rc, err);
    gsnp::matrix_gen(blocks, g, *p, pm);
    return GSNP_OK;
}

}  // extern "C"
