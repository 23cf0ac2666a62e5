// engine.h -- device buffers and the calling context behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <functional>
#include <memory>
#include <utility>
#include <new>
#include <string>
#include <vector>

#include "../../include/gsnp_abi.h"
#include "cols.h"
#include "kernels.h"

namespace gsnp {

int fail(int code, const std::string& msg);
// Runs an entry point's body; a C++ exception never crosses the C ABI: host
// allocation failures become GSNP_E_NOMEM, anything else GSNP_E_ARG.
template <class F>
int guarded(F body)
{
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(GSNP_E_NOMEM, "out of host memory");
    } catch (const std::exception& e) {
        return fail(GSNP_E_ARG, e.what());
    }
}
struct Ctx;
// calibration bins: (q - q_min) << 12 | cycle << 4 | ref << 2 | obs
constexpr uint32_t kCalibBins = 64u << 12;
int dict_encode_device(Ctx& c, const int32_t* d_in, uint64_t n, uint64_t* n_words, uint64_t* n_dict);

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void swap(DevBuf& o)
    {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
    hipError_t alloc(size_t count)
    {
        release();
        if (!count) count = 1;
        n = count;
        return hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T));
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct SiteBufs {
    DevBuf<uint32_t> rec;
    DevBuf<int32_t> nrec, depth, repeat;
    DevBuf<int4> cu, qs, ca;
    void reset()
    {
        rec.release();
        nrec.release();
        depth.release();
        repeat.release();
        cu.release();
        qs.release();
        ca.release();
    }
};

// Stages timed with HIP events on the context stream when profiling is on.
enum Stage : int {
    kStSiteBuild = 0,
    kStSort,
    kStLikelihood,
    kStPosterior,
    kStRank,
    kStFormat,
    kStRecycle,
    kStFused,
    kStAdjoint,
    kStColumns,
    kStCalib,
    kStBucket,
    kNumStages
};

struct StageTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[kNumStages];
    hipEvent_t open[kNumStages] = {};
    double ms[kNumStages] = {};
    uint64_t count[kNumStages] = {};
    hipEvent_t take();
    void begin(int st, hipStream_t s);
    void end(int st, hipStream_t s);
    int collect();
    ~StageTimer();
};

// An instantiated graph of a captured launch sequence and the key (shape and
// buffer generation) it was captured for.
struct GraphSlot {
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> key;
    GraphSlot() = default;
    GraphSlot(const GraphSlot&) = delete;
    GraphSlot& operator=(const GraphSlot&) = delete;
    void reset()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    ~GraphSlot() { reset(); }
};

// Device buffers of the -C column encoder, sized for `win` windows of
// GSNP_WIN sites (segment w of each per-window array starts at w*GSNP_WIN).
struct ColBufs {
    uint32_t win = 0;
    DevBuf<uint32_t> cnt, boff, tot;  // per 256-site block and per window: 16 counters
    DevBuf<int32_t> id4, id6, idn, nidx, x12i, x12v, x13i, x13v, cid, id15, id17;
    DevBuf<uint8_t> v4, v6;
    DevBuf<uint32_t> code3, code10;  // 2-bit packed columns 3 and 10
    DevBuf<double> val15;
    DevBuf<uint32_t> rstart;  // run starts of the six run-length columns
    DevBuf<uint32_t> in32;    // the 15 dictionary-coded arrays (32-bit ones)
    DevBuf<uint64_t> in64;    // run values of column 16 (f64 bits)
    DevBuf<uint64_t> tab, dkeys;
    DevBuf<uint32_t> trank, dslot, eslot, words, dict32, nd, nw;
    DevBuf<uint64_t> dict64, foff, ftot;
    // integer dictionaries spread over workgroups: presence bits (zero
    // between passes), their copy and word prefix, and per array the largest
    // value, an out-of-range flag and a done flag
    DevBuf<uint32_t> gbits, rbits, wpre, gstate;
    // the copy numbers' split path: a hash set per window (empty between
    // passes), each slot's rank, the distinct values and their slots
    DevBuf<uint64_t> htab, hkeys;
    DevBuf<uint32_t> hrank, hslot;
    // the five file streams of a pass, double-buffered: the copy of one pass
    // to pinned host memory runs on `copy` while the next pass is encoded
    DevBuf<uint8_t> stream[2];
    uint64_t stream_cap[2] = {0, 0};
    uint8_t* host[2] = {nullptr, nullptr};
    uint64_t host_cap[2] = {0, 0};
    hipStream_t copy = nullptr;
    hipEvent_t ev_written = nullptr, ev_copied[2] = {nullptr, nullptr};
    int cur = 0;            // the buffer the next pass takes
    bool have_prev = false;  // prev: the last pass queued, not yet handed over
    ColumnPending prev;
    ColBufs() = default;
    ColBufs(const ColBufs&) = delete;
    ColBufs& operator=(const ColBufs&) = delete;
    ~ColBufs()
    {
        if (copy) (void)hipStreamSynchronize(copy);
        for (uint8_t* h : host)
            if (h) (void)hipHostFree(h);
        for (hipEvent_t e : {ev_written, ev_copied[0], ev_copied[1]})
            if (e) (void)hipEventDestroy(e);
        if (copy) (void)hipStreamDestroy(copy);
    }
};

// The device inputs of one chromosome: reference codes, known SNPs and the
// packed read store.  The context holds the selected chromosome's in its own
// fields; a job that keeps every chromosome resident parks the others in
// ChromStates and swaps them in (Ctx::swap_chrom) without an upload.
struct ChromState {
    uint64_t chrom_id = 0;
    std::string chr_name;
    uint32_t chr_len = 0;
    DevBuf<uint8_t> ref;
    DevBuf<uint32_t> snp_pos, snp_first;
    DevBuf<double> snp_lp;
    uint32_t n_snp = 0;
    DevBuf<char> d_name;
    uint64_t n_reads = 0;
    uint32_t pitch = 0;
    DevBuf<int32_t> rpos;
    DevBuf<uint32_t> rmeta;
    DevBuf<uint8_t> rseq;
    std::vector<int32_t> h_rpos;
    std::vector<gsnp_snp> h_snps;
    std::vector<uint8_t> h_snp_ref;
};

// .cns rows on their way to host memory: the rows of a span are formatted
// into one of two device buffers while the copy of the previous span's rows
// to pinned host memory runs on a stream of its own.
struct TextPipe {
    // spans whose rows may be in flight at once: while one span is computed,
    // the copies of up to three earlier ones stay queued, so that a short span
    // followed by a long one leaves no gap on the copy engine
    static constexpr int kBufs = 4;
    DevBuf<char> dev[kBufs];
    char* host[kBufs] = {};
    uint64_t host_cap[kBufs] = {};
    uint64_t len[kBufs] = {};
    bool recorded[kBufs] = {};  // ev_done[i] has been recorded
    int cur = 0;                // buffer the next span is formatted into
    int pending = 0;            // spans queued and not yet handed on, the oldest in buffer cur - pending
    hipStream_t copy = nullptr;
    hipEvent_t ev_done[kBufs] = {};  // [i]: host[i] holds its span's rows
    hipEvent_t ev_fmt = nullptr;
    TextPipe() = default;
    TextPipe(const TextPipe&) = delete;
    TextPipe& operator=(const TextPipe&) = delete;
    ~TextPipe()
    {
        if (copy) (void)hipStreamSynchronize(copy);
        for (char* h : host)
            if (h) (void)hipHostFree(h);
        for (hipEvent_t e : ev_done)
            if (e) (void)hipEventDestroy(e);
        if (ev_fmt) (void)hipEventDestroy(ev_fmt);
        if (copy) (void)hipStreamDestroy(copy);
    }
};

struct Ctx {
    gsnp_params p{};
    int device = 0;  // HIP device current when the context was created
    hipStream_t stream = nullptr;      // where the context's work is queued
    hipStream_t own_stream = nullptr;  // the stream the context created (stream unless a caller's is set)
    StageTimer timer;
    // tables
    DevBuf<double> score, lprior, prank;
    DevBuf<uint8_t> qadj;
    int td = 1, gd = 1;
    std::vector<double> h_prank, h_pprior;
    // chromosome (the fields of ChromState; chrom_id names the uploaded
    // inputs in the keys of captured graphs)
    uint64_t chrom_id = 0;
    std::string chr_name;
    uint32_t chr_len = 0;
    DevBuf<uint8_t> ref;
    DevBuf<uint32_t> snp_pos, snp_first;
    DevBuf<double> snp_lp;
    uint32_t n_snp = 0;
    DevBuf<char> d_name;
    // reads
    uint64_t n_reads = 0;
    uint32_t pitch = 0;
    DevBuf<int32_t> rpos;
    DevBuf<uint32_t> rmeta;
    DevBuf<uint8_t> rseq;
    std::vector<int32_t> h_rpos;
    // sites of the current window (or batched span)
    uint32_t ws = 0;
    uint32_t win_cap = 0;
    bool built = false;
    SiteBufs sites;
    DevBuf<double> lk;
    DevBuf<uint8_t> base1, base2, type1;
    DevBuf<int32_t> qcns;
    DevBuf<double> rank;
    DevBuf<uint32_t> lists, counts;
    DevBuf<RankJob> jobs;
    DevBuf<uint32_t> job_count;
    DevBuf<FixJob> fixes;
    DevBuf<uint32_t> fix_count;
    uint32_t fix_offset = 0;
    uint32_t last_fixups = 0;
    // sites finished on the host since the last reset: every one, and those
    // whose rank-sum test took the normal approximation (n1 + n2 >= 64)
    uint64_t fixups_total = 0, fixups_normal = 0;
    // text
    DevBuf<uint32_t> rowlen;
    DevBuf<uint64_t> blocksum, total;
    DevBuf<uint64_t> ranges;  // read interval per 64-site block (site build)
    DevBuf<uint32_t> gidx, gcount;  // rows whose %#g columns need the exact routine
    DevBuf<GText> gtext;
    bool keep_lk = false;     // batched form: also store the per-site log-likelihoods
    FusedTune tune;           // launch shape of the fused stage (gsnp_autotune)
    // batched span and its formatting replay as graphs while shape and buffers
    // stay the same; buf_gen counts buffer (re)allocations and chromosome or
    // read changes, each made after a stream sync, and drops the cached graphs
    bool use_graphs = true;
    uint64_t buf_gen = 0, graphs_gen = 0;
    std::vector<std::unique_ptr<GraphSlot>> graphs;  // captured spans, at most kMaxGraphs
    GraphSlot& graph_slot(const std::vector<uint64_t>& key);
    DevBuf<char> text;  // rows of the per-window path and of spans formatted for gsnp_fetch_text
    uint64_t text_cap = 0, text_len = 0;
    TextPipe pipe;      // rows of batched spans copied to host memory (gsnp_job_call, gsnp_run)
    DevBuf<uint32_t> err;
    DevBuf<unsigned long long> stat;  // [0] records called
    DevBuf<int4> head_qs;  // q_sum of window-boundary sites before their own reads (batched span)
    // known SNPs of the chromosome (host copies, for the adjoint's prior rows)
    std::vector<gsnp_snp> h_snps;
    std::vector<uint8_t> h_snp_ref;
    // adjoint: records of a prepared span in site order, their score-row keys,
    // their place in key order and the reduction pieces
    uint32_t adj_start = 0, adj_n = 0, adj_npieces = 0;
    uint64_t adj_nrec = 0;
    DevBuf<uint32_t> adj_cnt, adj_prow, adj_key, adj_skey, adj_perm, adj_tmp, adj_pstart, adj_kpiece;
    DevBuf<uint32_t> adj_rsite;  // per record (record order): its site
    DevBuf<uint32_t> adj_ssite;  // per record in key order: its site
    DevBuf<uint32_t> adj_hist, adj_mask;
    DevBuf<uint64_t> adj_off, adj_pidx, adj_hoff, adj_bsum;
    DevBuf<double> adj_T, adj_w, adj_ll, adj_dh, adj_psum, adj_part, adj_out, adj_pm, adj_grad, adj_lp;
    // device dictionary encoder scratch
    DevBuf<int32_t> dict_mm, dict_vals;
    DevBuf<uint32_t> dict_bits, dict_wpre, dict_out;
    DevBuf<uint64_t> dict_bsum;
    ColBufs cols;  // -C column encoder
    DevBuf<unsigned long long> calib;  // calibration sub-histograms

    int set_matrix(const double* p_matrix);
    void swap_chrom(ChromState& s);
    // rows of the span just called (gsnp_batched_call without text) formatted
    // into the pipe's device buffer and their copy to pinned host memory queued
    // on the copy stream; the rows of the span TextPipe::kBufs spans back are
    // handed to consume(ptr, len) first, once landed, when the pipe is full.
    // finish_text hands on the spans still queued, in order.
    using TextSink = std::function<int(const char*, uint64_t)>;
    int text_to_host(uint32_t start, uint32_t n, uint64_t* nbytes, const TextSink& consume);
    int finish_text(const TextSink& consume);
    // drops rows still in the pipe (a pass that ended early) and waits for its copies
    int reset_text();
    // the rows of the last gsnp_window_text copied to pinned host memory and
    // handed to consume, synchronously (the per-window path)
    int window_text_to_host(const TextSink& consume);
    int alloc_window(uint32_t n_sites);
    SiteArrays site_view(uint32_t offset);
    Results result_view(uint32_t offset);
    Tables tables() const;
    RefView ref_view() const;
    ReadView read_view() const;
    int host_fixups();
    int check_err();
    int format_text(uint32_t offset, uint32_t ws_abs, uint32_t n_sites, uint32_t pos_base, uint64_t* nbytes,
                    bool graph = false, char* out = nullptr);
    int span_build(uint32_t start, uint32_t n);
    int span_prepare(uint32_t n);
    void span_enqueue(uint32_t start, uint32_t n);
    template <class F>
    int launch_captured(GraphSlot& slot, const std::vector<uint64_t>& key, F enqueue);
    FusedArgs fused_args(uint32_t start);
};

int create_ctx(const gsnp_params* p, const double* p_matrix, gsnp_ctx** out);
// device address of the fixed working array of gsnp_sort_likelihood_noalloc
int noalloc_scratch(int device, uint32_t** out);

}  // namespace gsnp

// the opaque handle of the C ABI is the context itself
struct gsnp_ctx : public gsnp::Ctx {};
