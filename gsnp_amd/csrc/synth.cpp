// synth.cpp -- seeded synthetic SOAPsnp inputs (FASTA, SOAP alignments, known-SNP
// table) for the five benchmark configurations, generated on threads.
//
// PRNG: xoshiro256** (Blackman & Vigna 2018) seeded through splitmix64.  Every
// piece of a data set -- a 64 kb chunk of a contig's reference and haplotypes,
// the reads started in one depth block, a contig's known SNPs -- draws from its
// own stream, keyed by (config, seed, contig, piece), so pieces are generated
// in parallel and a (config, seed, scale) triple still names one byte-identical
// input set, however many threads generate it.
//
// Configurations (genome length multiplied by `scale`, read count follows):
//   1  100 kb, 20x single-end 100 bp, q ~ N(30,7) on [2,40]
//   2  10 Mb, 30x 100 bp pairs (insert 500+-50), 3% hit=2, q mean 36->24 by
//      cycle (sd 6), 1% known-SNP prior file
//   3  24 contigs log-uniform 1-20 Mb, depth per 10 kb ~ Gamma(4, 7.5), 1%
//      blocks at depth 200-255 with hits 1-3, 2% uncovered
//   4  2 Mb 30x, errors from q - delta, delta = 3[q>30] + 2(c-1)/99
//   5  248 Mb 30x 150 bp pairs (insert 400+-40), q decaying by cycle
//   6  stress: 3 contigs (560 kb, 300 kb, 90 kb) with N runs, depth blocks from
//      0 to 300x (past the 255 record cap), hits 0-3, N and low-quality read
//      bases, insertion/deletion edits, 75 bp reads, a known-SNP file
#include "synth.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

namespace gsnp {

namespace {

struct Xoshiro256ss {
    uint64_t s[4];
    static uint64_t splitmix64(uint64_t& x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit Xoshiro256ss(uint64_t seed)
    {
        for (auto& v : s) v = splitmix64(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
    double normal()
    {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    double gamma_int(int k, double theta)
    {
        double x = 0;
        for (int i = 0; i < k; i++) x += -std::log(1.0 - uniform());
        return x * theta;
    }
};

// the stream of one piece of a data set
Xoshiro256ss stream(uint64_t seed, int config, uint64_t contig, uint64_t kind, uint64_t piece)
{
    uint64_t x = seed * 0x100000001b3ULL + (uint64_t)config;
    uint64_t h = Xoshiro256ss::splitmix64(x);
    for (uint64_t v : {contig, kind, piece}) {
        x = h ^ (v + 0x632be59bd9b4e019ULL);
        h = Xoshiro256ss::splitmix64(x);
    }
    return Xoshiro256ss(h);
}

enum : uint64_t { kKindContig = 0, kKindRef = 1, kKindReads = 2, kKindSnp = 3 };
constexpr uint64_t kRefChunk = 65536;

// base codes A0 C1 T2 G3 (the reference's binary order); transition partner is code^3
const char kBase[4] = {'A', 'C', 'T', 'G'};

struct Cfg {
    int id = 1;
    std::vector<uint64_t> lens;
    int read_len = 100;
    bool paired = false;
    double insert_mean = 500, insert_sd = 50;
    double depth = 20;
    double frac_hit2 = 0;
    int qual_model = 0;  // 0 iid N(30,7); 1 cycle mean 36->24 sd 6; 2 decaying by cycle
    bool miscal = false; // config 4 error model
    bool gamma_depth = false;
    bool dbsnp = false;
    bool stress = false;
    uint64_t block = 10000;  // sites per depth block (each block draws its reads from its own stream)
};

Cfg config_for(int id, double scale, uint64_t seed)
{
    Cfg c;
    c.id = id;
    auto sc = [&](double v) { return (uint64_t)std::max(1000.0, std::floor(v * scale)); };
    switch (id) {
    case 1:
        c.lens = {sc(100000)};
        c.depth = 20;
        break;
    case 2:
        c.lens = {sc(10000000)};
        c.depth = 30;
        c.paired = true;
        c.frac_hit2 = 0.03;
        c.qual_model = 1;
        c.dbsnp = true;
        break;
    case 3: {
        Xoshiro256ss g = stream(seed, id, 0, kKindContig, 0);
        for (int i = 0; i < 24; i++) c.lens.push_back(sc(std::exp(std::log(1e6) + g.uniform() * std::log(20.0))));
        c.paired = true;
        c.gamma_depth = true;
        break;
    }
    case 4:
        c.lens = {sc(2000000)};
        c.depth = 30;
        c.miscal = true;
        c.qual_model = 0;
        break;
    case 5:
        c.lens = {sc(248000000)};
        c.depth = 30;
        c.read_len = 150;
        c.paired = true;
        c.insert_mean = 400;
        c.insert_sd = 40;
        c.qual_model = 2;
        break;
    case 6:
        c.lens = {sc(560000), sc(300000), sc(90000)};
        c.read_len = 75;
        c.depth = 25;
        c.stress = true;
        c.dbsnp = true;
        c.qual_model = 1;
        c.block = 5000;
        break;
    default:
        c.lens = {sc(100000)};
    }
    return c;
}

// Quality draw per cycle: floor(m + sd * z + 0.5) for a standard normal z,
// redrawn until it lies in [2, 40], as a table of the 39 cumulative
// probabilities (Phi by erfc), so one uniform draws one quality.
struct QualTable {
    std::vector<double> cdf;  // [cycle][39]
    QualTable(const Cfg& c)
    {
        cdf.resize((size_t)c.read_len * 39);
        for (int cy = 0; cy < c.read_len; cy++) {
            double m, sd;
            if (c.qual_model == 0) {
                m = 30, sd = 7;
            } else if (c.qual_model == 1) {
                m = 36.0 - 12.0 * cy / 99.0;
                sd = 6;
            } else {
                m = 38.0 - 14.0 * cy / (c.read_len - 1.0);
                sd = 5;
            }
            auto phi = [&](double x) { return 0.5 * std::erfc(-(x - m) / (sd * std::sqrt(2.0))); };
            double acc = 0, tot = phi(40.5) - phi(1.5);
            for (int q = 2; q <= 40; q++) {
                acc += phi(q + 0.5) - phi(q - 0.5);
                cdf[(size_t)cy * 39 + (q - 2)] = acc / tot;
            }
            cdf[(size_t)cy * 39 + 38] = 1.0;
        }
    }
    int draw(Xoshiro256ss& g, int cycle) const
    {
        const double u = g.uniform();
        const double* t = &cdf[(size_t)cycle * 39];
        int lo = 0, hi = 38;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (u < t[mid])
                hi = mid;
            else
                lo = mid + 1;
        }
        return lo + 2;
    }
};

uint8_t mutate(Xoshiro256ss& g, uint8_t b)
{
    // Ti:Tv 2:1 -> transition partner with probability 2/3, each transversion 1/6
    double u = g.uniform();
    if (u < 2.0 / 3.0) return b ^ 3;
    uint8_t a = (uint8_t)(b ^ 1), c = (uint8_t)(b ^ 2);
    return u < 5.0 / 6.0 ? a : c;
}

// Runs f(i) for i in [0, n) on up to `threads` threads.
template <class F>
void parallel_for(uint64_t n, int threads, F f)
{
    std::atomic<uint64_t> next{0};
    auto work = [&] {
        for (uint64_t i; (i = next.fetch_add(1)) < n;) f(i);
    };
    std::vector<std::thread> th;
    for (int t = 1; t < threads && (uint64_t)t < n; t++) {
        try {
            th.emplace_back(work);
        } catch (const std::system_error&) {
            break;
        }
    }
    work();
    for (auto& x : th) x.join();
}

struct ContigWork {
    std::vector<uint8_t> hap[2];
};

// The reads whose fragment starts fall in one depth block.
struct BlockReads {
    std::vector<SynthRead> r;
    std::vector<char> seq, qual;
};

struct Generator {
    const Cfg& c;
    uint64_t seed;
    const QualTable& qt;
    std::vector<double> perr;  // [cycle][64] error probability 10^(-(q - delta)/10)

    Generator(const Cfg& cfg, uint64_t sd, const QualTable& q) : c(cfg), seed(sd), qt(q)
    {
        perr.assign((size_t)c.read_len * 64, 0.0);
        for (int cy = 0; cy < c.read_len; cy++)
            for (int q = 0; q < 64; q++) {
                double qe = q;
                if (c.miscal) qe = q - (3.0 * (q > 30) + 2.0 * cy / 99.0);
                perr[(size_t)cy * 64 + q] = std::pow(10.0, -qe / 10.0);
            }
    }

    void emit(Xoshiro256ss& g, const SynthContig& ct, const ContigWork& w, int hap, int64_t pos, uint8_t strand, char ab,
              uint8_t hit, uint32_t id, BlockReads& out) const
    {
        const int L = c.read_len;
        SynthRead r{};
        r.pos = (int32_t)pos;
        r.strand = strand;
        r.hit = hit;
        r.ab = ab;
        r.id = id;
        const size_t at = out.seq.size();
        out.seq.resize(at + L);
        out.qual.resize(at + L);
        char* s = &out.seq[at];
        char* qq = &out.qual[at];
        for (int k = 0; k < L; k++) {
            const int cycle = strand ? L - 1 - k : k;
            const int q = qt.draw(g, cycle);
            uint8_t b = w.hap[hap][pos + k];
            if (b == 7) b = (uint8_t)g.below(4);  // reads over reference N carry real bases
            if (g.uniform() < perr[(size_t)cycle * 64 + q]) b = (uint8_t)((b + 1 + g.below(3)) & 3);
            s[k] = kBase[b];
            qq[k] = (char)(64 + q);
            if (c.stress) {
                const double u = g.uniform();
                if (u < 0.01)
                    s[k] = 'N';
                else if (u < 0.02)
                    qq[k] = (char)(64 - 1 - (int)g.below(3));  // below q_min: not observed
            }
        }
        if (c.stress && g.uniform() < 0.01) {  // an insertion or deletion edit of 1-3 bases
            r.indel = (int16_t)(g.below(2) ? 100 + 1 + (int)g.below(3) : 200 + 1 + (int)g.below(3));
            r.indel_pos = (int16_t)(5 + (int)g.below(L - 10));
        }
        (void)ct;
        out.r.push_back(r);
    }

    // reads whose fragments start in [lo, hi) at the given depth
    void span(Xoshiro256ss& g, const SynthContig& ct, const ContigWork& w, uint64_t lo, uint64_t hi, double depth,
              bool repeat_like, uint32_t id0, BlockReads& out) const
    {
        const int L = c.read_len;
        const uint64_t len = ct.ref.size();
        if (len < (uint64_t)L + 2 || depth <= 0) return;
        double nreads = depth * (double)(hi - lo) / L;
        uint64_t n = (uint64_t)nreads;
        if (g.uniform() < nreads - (double)n) n++;
        if (c.paired) n /= 2;
        uint32_t id = id0;
        for (uint64_t i = 0; i < n; i++) {
            int hap = (int)g.below(2);
            uint8_t hit = 1;
            if (repeat_like)
                hit = (uint8_t)(1 + g.below(3));
            else if (c.frac_hit2 > 0 && g.uniform() < c.frac_hit2)
                hit = 2;
            else if (c.stress && g.uniform() < 0.02)
                hit = (uint8_t)(g.below(2) ? 0 : 3);
            if (!c.paired) {
                int64_t pos = (int64_t)(lo + g.below(hi - lo));
                if (pos + L + 1 > (int64_t)len) pos = (int64_t)len - L - 1;
                emit(g, ct, w, hap, pos, (uint8_t)g.below(2), 'a', hit, id++, out);
            } else {
                int64_t ins = (int64_t)std::llround(c.insert_mean + c.insert_sd * g.normal());
                if (ins < L) ins = L;
                int64_t f = (int64_t)(lo + g.below(hi - lo));
                if (f + ins + 1 > (int64_t)len) f = (int64_t)len - ins - 1;
                if (f < 0) continue;
                emit(g, ct, w, hap, f, 0, 'a', hit, id, out);
                emit(g, ct, w, hap, f + ins - L, 1, 'b', hit, id, out);
                id++;
            }
        }
    }

    void block(uint64_t ci, uint64_t bi, const SynthContig& ct, const ContigWork& w, BlockReads& out) const
    {
        Xoshiro256ss g = stream(seed, c.id, ci, kKindReads, bi);
        const uint64_t lo = bi * c.block, hi = std::min<uint64_t>(ct.ref.size(), lo + c.block);
        // read ids: a block's pairs/reads are numbered from its own base, unique in the contig
        const uint32_t id0 = (uint32_t)(bi << 16);
        if (c.stress) {
            const double u = g.uniform();
            const double d = u < 0.05 ? 0 : u < 0.12 ? 260 + (double)g.below(40) : u < 0.2 ? 80 + (double)g.below(60)
                                                                                            : 5 + 30 * g.uniform();
            span(g, ct, w, lo, hi, d, u >= 0.05 && u < 0.09, id0, out);
        } else if (c.gamma_depth) {
            const double u = g.uniform();
            if (u < 0.02) return;  // uncovered block
            if (u < 0.03)
                span(g, ct, w, lo, hi, 200 + (double)g.below(56), true, id0, out);  // repeat-like block
            else
                span(g, ct, w, lo, hi, g.gamma_int(4, 7.5), false, id0, out);
        } else {
            span(g, ct, w, lo, hi, c.depth, false, id0, out);
        }
    }
};

void make_contig(const Cfg& c, uint64_t seed, uint64_t ci, SynthContig& ct, ContigWork& w, int threads)
{
    const uint64_t len = ct.ref.size();
    w.hap[0].resize(len);
    w.hap[1].resize(len);
    const uint64_t nchunk = (len + kRefChunk - 1) / kRefChunk;
    parallel_for(nchunk, threads, [&](uint64_t k) {
        Xoshiro256ss g = stream(seed, c.id, ci, kKindRef, k);
        const uint64_t lo = k * kRefChunk, hi = std::min(len, lo + kRefChunk);
        for (uint64_t i = lo; i < hi; i++) {
            const double u = g.uniform();  // P(A)=P(T)=0.295, P(C)=P(G)=0.205
            ct.ref[i] = u < 0.295 ? 0 : (u < 0.5 ? 1 : (u < 0.795 ? 2 : 3));
        }
        for (uint64_t i = lo; i < hi; i++) {
            w.hap[0][i] = w.hap[1][i] = ct.ref[i];
            const double u = g.uniform();
            if (u < 1e-3) {  // het
                w.hap[g.below(2)][i] = mutate(g, ct.ref[i]);
            } else if (u < 1.5e-3) {  // hom alt
                const uint8_t b = mutate(g, ct.ref[i]);
                w.hap[0][i] = w.hap[1][i] = b;
            }
        }
    });
    if (c.stress) {  // a few N runs in the reference (code 7), carried into the haplotypes
        Xoshiro256ss g = stream(seed, c.id, ci, kKindContig, 1);
        for (int k = 0; k < 4; k++) {
            const uint64_t at = g.below(len - 2000), run = 50 + g.below(1500);
            for (uint64_t i = at; i < at + run; i++) ct.ref[i] = w.hap[0][i] = w.hap[1][i] = 7;
        }
    }
}

// 1% of positions: half of the true variant sites plus random others; ref +
// one alt allele, alt frequency U(0.05, 0.5); values at the six decimals the
// table is written with
void make_snps(const Cfg& c, uint64_t seed, uint64_t ci, SynthContig& ct, const ContigWork& w)
{
    Xoshiro256ss g = stream(seed, c.id, ci, kKindSnp, 0);
    const uint64_t len = ct.ref.size(), want = len / 100;
    std::vector<uint64_t> pos;
    for (uint64_t i = 0; i < len && pos.size() < want / 2; i++)
        if ((w.hap[0][i] != ct.ref[i] || w.hap[1][i] != ct.ref[i]) && g.uniform() < 0.5) pos.push_back(i);
    while (pos.size() < want) pos.push_back(g.below(len));
    std::sort(pos.begin(), pos.end());
    pos.erase(std::unique(pos.begin(), pos.end()), pos.end());
    for (uint64_t p : pos) {
        const uint8_t r0 = ct.ref[p];
        if (r0 > 3) continue;  // no known SNPs on reference N
        const uint8_t alt = w.hap[0][p] != r0 ? w.hap[0][p] : (w.hap[1][p] != r0 ? w.hap[1][p] : mutate(g, r0));
        SynthSnp s{};
        s.pos = (uint32_t)p;
        const double af = 0.05 + 0.45 * g.uniform();
        double f[4] = {0, 0, 0, 0};
        f[alt] = af;
        f[r0] = 1.0 - af;
        for (int k = 0; k < 4; k++) {
            char b[32];
            std::snprintf(b, sizeof b, "%.6f", f[k]);
            s.freq[k] = std::strtod(b, nullptr);
        }
        s.hapmap = g.uniform() < 0.1;
        s.validated = g.uniform() < 0.5;
        s.indel = g.uniform() < 0.02;
        ct.snps.push_back(s);
    }
}

}  // namespace

int synth_generate(int config, uint64_t seed, double scale, const std::vector<int>* only, int threads, SynthSet& out)
{
    const Cfg c = config_for(config, scale, seed);
    const QualTable qt(c);
    const Generator gen(c, seed, qt);
    if (threads < 1) threads = 1;
    out.read_len = c.read_len;
    out.contigs.clear();
    const int L = c.read_len;
    for (size_t ci = 0; ci < c.lens.size(); ci++) {
        if (only && std::find(only->begin(), only->end(), (int)ci) == only->end()) continue;
        out.contigs.emplace_back();
        SynthContig& ct = out.contigs.back();
        ct.name = c.lens.size() == 1 ? std::string("chr1") : ("chr" + std::to_string(ci + 1));
        ct.ref.resize(c.lens[ci]);
        ContigWork w;
        make_contig(c, seed, ci, ct, w, threads);
        const uint64_t nb = (ct.ref.size() + c.block - 1) / c.block;
        std::vector<BlockReads> blocks(nb);
        parallel_for(nb, threads, [&](uint64_t b) { gen.block(ci, b, ct, w, blocks[b]); });
        if (c.dbsnp) make_snps(c, seed, ci, ct, w);
        // position order, stable over block order (a counting sort by start)
        uint64_t n = 0;
        for (const BlockReads& b : blocks) n += b.r.size();
        std::vector<uint32_t> first(ct.ref.size() + 1, 0);
        for (const BlockReads& b : blocks)
            for (const SynthRead& r : b.r) first[r.pos + 1]++;
        for (size_t i = 1; i < first.size(); i++) first[i] += first[i - 1];
        ct.reads.resize(n);
        ct.seq.resize(n * (size_t)L);
        ct.qual.resize(n * (size_t)L);
        std::vector<uint64_t> boff(nb + 1, 0);  // first read of each block
        for (uint64_t b = 0; b < nb; b++) boff[b + 1] = boff[b] + blocks[b].r.size();
        // destinations in block order (serial: keeps the sort stable), copies on threads
        std::vector<uint32_t> dest(n);
        for (uint64_t b = 0; b < nb; b++)
            for (size_t k = 0; k < blocks[b].r.size(); k++) dest[boff[b] + k] = first[blocks[b].r[k].pos]++;
        parallel_for(nb, threads, [&](uint64_t b) {
            BlockReads& br = blocks[b];
            for (size_t k = 0; k < br.r.size(); k++) {
                const uint32_t d = dest[boff[b] + k];
                ct.reads[d] = br.r[k];
                std::memcpy(&ct.seq[(size_t)d * L], &br.seq[k * L], L);
                std::memcpy(&ct.qual[(size_t)d * L], &br.qual[k * L], L);
            }
            BlockReads().r.swap(br.r);
            std::vector<char>().swap(br.seq);
            std::vector<char>().swap(br.qual);
        });
    }
    return 0;
}

int synth_contig_count(int config, uint64_t seed, double scale, std::vector<uint64_t>* lens)
{
    const Cfg c = config_for(config, scale, seed);
    if (lens) *lens = c.lens;
    return (int)c.lens.size();
}

// FASTA, SOAP and (configs 2 and 6) known-SNP files of a generated set.
long long synth_write_files(const SynthSet& s, const std::string& prefix, int threads)
{
    const int L = s.read_len;
    FILE* fa = std::fopen((prefix + ".fa").c_str(), "w");
    FILE* so = std::fopen((prefix + ".soap").c_str(), "w");
    bool any_snp = false;
    for (const SynthContig& ct : s.contigs) any_snp = any_snp || !ct.snps.empty();
    FILE* sn = any_snp ? std::fopen((prefix + ".snp").c_str(), "w") : nullptr;
    bool ok = fa && so && (!any_snp || sn);
    long long nlines = 0;
    std::vector<char> line(4096);
    for (const SynthContig& ct : s.contigs) {
        if (!ok) break;
        ok = std::fprintf(fa, ">%s\n", ct.name.c_str()) > 0;
        for (uint64_t i = 0; ok && i < ct.ref.size(); i += 60) {
            const uint64_t m = std::min<uint64_t>(60, ct.ref.size() - i);
            for (uint64_t k = 0; k < m; k++) line[k] = ct.ref[i + k] == 7 ? 'N' : kBase[ct.ref[i + k]];
            line[m] = '\n';
            ok = std::fwrite(line.data(), 1, m + 1, fa) == m + 1;
        }
        // SOAP lines, formatted on threads a chunk at a time and written in order
        const uint64_t n = ct.reads.size(), chunk = 1 << 16;
        const uint64_t nch = (n + chunk - 1) / chunk;
        for (uint64_t c0 = 0; ok && c0 < nch; c0 += 64) {
            const uint64_t c1 = std::min(nch, c0 + 64);
            std::vector<std::string> text(c1 - c0);
            parallel_for(c1 - c0, threads, [&](uint64_t k) {
                std::string& t = text[k];
                char buf[1024];
                std::string rs(L, ' '), qs(L, ' ');
                const uint64_t lo = (c0 + k) * chunk, hi = std::min(n, lo + chunk);
                for (uint64_t i = lo; i < hi; i++) {
                    const SynthRead& r = ct.reads[i];
                    rs.assign(&ct.seq[i * L], L);
                    qs.assign(&ct.qual[i * L], L);
                    int mm = 0;
                    for (int k2 = 0; k2 < L; k2++) {
                        const uint8_t rc = ct.ref[r.pos + k2];
                        mm += rs[k2] != (rc == 7 ? 'N' : kBase[rc]);
                    }
                    if (mm > 99) mm = 99;
                    int m;
                    if (r.indel)
                        m = std::snprintf(buf, sizeof buf, "r%u/%c\t%s\t%s\t%d\t%c\t%d\t%c\t%s\t%d\t%d\t%d\n", r.id,
                                          r.ab == 'a' ? '1' : '2', rs.c_str(), qs.c_str(), (int)r.hit, r.ab, L,
                                          r.strand ? '-' : '+', ct.name.c_str(), r.pos + 1, (int)r.indel,
                                          (int)r.indel_pos);
                    else
                        m = std::snprintf(buf, sizeof buf, "r%u/%c\t%s\t%s\t%d\t%c\t%d\t%c\t%s\t%d\t%d\n", r.id,
                                          r.ab == 'a' ? '1' : '2', rs.c_str(), qs.c_str(), (int)r.hit, r.ab, L,
                                          r.strand ? '-' : '+', ct.name.c_str(), r.pos + 1, mm);
                    t.append(buf, (size_t)m);
                }
            });
            for (const std::string& t : text)
                if (ok && !t.empty()) ok = std::fwrite(t.data(), 1, t.size(), so) == t.size();
        }
        nlines += (long long)n;
        for (const SynthSnp& p : ct.snps) {
            if (!ok) break;
            ok = std::fprintf(sn, "%s\t%llu\t%d\t%d\t%d\t%.6f\t%.6f\t%.6f\t%.6f\trs%llu\n", ct.name.c_str(),
                              (unsigned long long)(p.pos + 1), (int)p.hapmap, (int)p.validated, (int)p.indel,
                              p.freq[0], p.freq[1], p.freq[2], p.freq[3], (unsigned long long)p.pos) > 0;
        }
    }
    for (FILE* f : {fa, so, sn})
        if (f) ok = std::fclose(f) == 0 && ok;
    return ok ? nlines : -1;
}

int synth_threads()
{
    const unsigned hw = std::thread::hardware_concurrency();
    int t = (int)(hw ? hw : 1);
    // a host share smaller than the machine (OMP_NUM_THREADS) caps it
    if (const char* omp = std::getenv("OMP_NUM_THREADS"))
        if (std::atoi(omp) > 0) t = std::min(t, std::atoi(omp));
    if (const char* env = std::getenv("GSNP_SYNTH_THREADS")) t = std::atoi(env);
    return std::max(1, std::min(t, 64));
}

}  // namespace gsnp

extern "C" {

// Writes <prefix>.fa, <prefix>.soap and (configs 2 and 6) <prefix>.snp.
// Returns the number of alignment lines written, or a negative value on error.
// The read length of the configuration is returned through *read_len.
long long gsnp_synth_write(int config, unsigned long long seed, double scale, const char* prefix, int* read_len)
{
    if (!prefix) return -1;
    try {
        gsnp::SynthSet s;
        const int t = gsnp::synth_threads();
        gsnp::synth_generate(config, seed, scale, nullptr, t, s);
        if (read_len) *read_len = s.read_len;
        return gsnp::synth_write_files(s, prefix, t);
    } catch (const std::exception&) {
        return -1;
    }
}

// The same files for the contigs listed in `contigs` only (indices in
// generation order): a sample of a configuration's data set.
long long gsnp_synth_write_contigs(int config, unsigned long long seed, double scale, const char* prefix,
                                   const int* contigs, int n_contigs, int* read_len)
{
    if (!prefix || (n_contigs > 0 && !contigs)) return -1;
    try {
        gsnp::SynthSet s;
        const int t = gsnp::synth_threads();
        const std::vector<int> only(contigs, contigs + (n_contigs > 0 ? n_contigs : 0));
        gsnp::synth_generate(config, seed, scale, &only, t, s);
        if (read_len) *read_len = s.read_len;
        return gsnp::synth_write_files(s, prefix, t);
    } catch (const std::exception&) {
        return -1;
    }
}

// Contig count of a configuration; the lengths of the first `cap` go to lens.
int gsnp_synth_contigs(int config, unsigned long long seed, double scale, uint64_t* lens, int cap)
{
    std::vector<uint64_t> v;
    const int n = gsnp::synth_contig_count(config, seed, scale, &v);
    for (int i = 0; i < n && i < cap && lens; i++) lens[i] = v[i];
    return n;
}

}  // extern "C"
