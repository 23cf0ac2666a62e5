// synth.cpp -- seeded synthetic SOAPsnp inputs (FASTA, SOAP alignments, known-SNP
// table) for the five benchmark configurations.
//
// PRNG: xoshiro256** (Blackman & Vigna 2018) seeded through splitmix64.  All
// draws go through this generator, so a (config, seed, scale) triple names one
// byte-identical input set.
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
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

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

// base codes A0 C1 T2 G3 (the reference's binary order); transition partner is code^3
const char kBase[4] = {'A', 'C', 'T', 'G'};

struct Contig {
    std::string name;
    std::vector<uint8_t> ref;     // base codes
    std::vector<uint8_t> hap[2];  // diploid haplotypes
};

struct Read {
    int indel;      // SOAP mismatch code > 100 (deletion) or > 200 (insertion), 0 if none
    int indel_pos;
    int32_t pos;  // 0-based
    uint8_t strand;
    uint8_t hit;
    char ab;
    uint32_t id;
    std::string seq, qual;
};

struct Cfg {
    int id = 1;
    int ncontig = 1;
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
};

int draw_q(Xoshiro256ss& g, const Cfg& c, int cycle)
{
    for (;;) {
        double m, sd;
        if (c.qual_model == 0) {
            m = 30, sd = 7;
        } else if (c.qual_model == 1) {
            m = 36.0 - 12.0 * cycle / 99.0;
            sd = 6;
        } else {
            m = 38.0 - 14.0 * cycle / (c.read_len - 1.0);
            sd = 5;
        }
        double v = std::floor(m + sd * g.normal() + 0.5);
        if (v >= 2 && v <= 40) return (int)v;
    }
}

uint8_t mutate(Xoshiro256ss& g, uint8_t b)
{
    // Ti:Tv 2:1 -> transition partner with probability 2/3, each transversion 1/6
    double u = g.uniform();
    if (u < 2.0 / 3.0) return b ^ 3;
    uint8_t a = (uint8_t)(b ^ 1), c = (uint8_t)(b ^ 2);
    return u < 5.0 / 6.0 ? a : c;
}

void make_contig(Xoshiro256ss& g, Contig& ct, uint64_t len)
{
    ct.ref.resize(len);
    for (uint64_t i = 0; i < len; i++) {
        double u = g.uniform();  // P(A)=P(T)=0.295, P(C)=P(G)=0.205
        ct.ref[i] = u < 0.295 ? 0 : (u < 0.5 ? 1 : (u < 0.795 ? 2 : 3));
    }
    ct.hap[0] = ct.ref;
    ct.hap[1] = ct.ref;
    for (uint64_t i = 0; i < len; i++) {
        double u = g.uniform();
        if (u < 1e-3) {  // het
            ct.hap[g.below(2)][i] = mutate(g, ct.ref[i]);
        } else if (u < 1.5e-3) {  // hom alt
            uint8_t b = mutate(g, ct.ref[i]);
            ct.hap[0][i] = ct.hap[1][i] = b;
        }
    }
}

void emit_read(Xoshiro256ss& g, const Cfg& c, const Contig& ct, int hap, int64_t pos, uint8_t strand, char ab,
               uint8_t hit, uint32_t id, std::vector<Read>& out)
{
    Read r;
    r.indel = 0;
    r.indel_pos = 0;
    r.pos = (int32_t)pos;
    r.strand = strand;
    r.hit = hit;
    r.ab = ab;
    r.id = id;
    const int L = c.read_len;
    r.seq.resize(L);
    r.qual.resize(L);
    // error probability 10^(-(q - delta)/10) per (cycle, q), delta = 0 unless miscalibrated
    static thread_local std::vector<double> perr;
    static thread_local int perr_key = -1;
    const int key = L * 2 + (c.miscal ? 1 : 0);
    if (perr_key != key) {
        perr.assign((size_t)L * 64, 0.0);
        for (int cy = 0; cy < L; cy++)
            for (int q = 0; q < 64; q++) {
                double qe = q;
                if (c.miscal) qe = q - (3.0 * (q > 30) + 2.0 * cy / 99.0);
                perr[(size_t)cy * 64 + q] = std::pow(10.0, -qe / 10.0);
            }
        perr_key = key;
    }
    for (int k = 0; k < L; k++) {
        int cycle = strand ? L - 1 - k : k;
        int q = draw_q(g, c, cycle);
        uint8_t b = ct.hap[hap][pos + k];
        if (b == 7) b = (uint8_t)g.below(4);  // reads over reference N carry real bases
        if (g.uniform() < perr[(size_t)cycle * 64 + q]) b = (uint8_t)((b + 1 + g.below(3)) & 3);
        r.seq[k] = kBase[b];
        r.qual[k] = (char)(64 + q);
        if (c.stress) {
            double u = g.uniform();
            if (u < 0.01) r.seq[k] = 'N';
            else if (u < 0.02) r.qual[k] = (char)(64 - 1 - (int)g.below(3));  // below q_min: not observed
        }
    }
    out.push_back(std::move(r));
}

// reads covering [lo, hi) at the given depth
void reads_for_span(Xoshiro256ss& g, const Cfg& c, const Contig& ct, uint64_t lo, uint64_t hi, double depth,
                    bool repeat_like, uint32_t& id, std::vector<Read>& out)
{
    const int L = c.read_len;
    const uint64_t len = ct.ref.size();
    if (len < (uint64_t)L + 2 || depth <= 0) return;
    double nreads = depth * (double)(hi - lo) / L;
    uint64_t n = (uint64_t)nreads;
    if (g.uniform() < nreads - (double)n) n++;
    if (c.paired) n /= 2;
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
            emit_read(g, c, ct, hap, pos, (uint8_t)g.below(2), 'a', hit, id++, out);
        } else {
            int64_t ins = (int64_t)std::llround(c.insert_mean + c.insert_sd * g.normal());
            if (ins < L) ins = L;
            int64_t f = (int64_t)(lo + g.below(hi - lo));
            if (f + ins + 1 > (int64_t)len) f = (int64_t)len - ins - 1;
            if (f < 0) continue;
            uint8_t hit2 = hit;
            emit_read(g, c, ct, hap, f, 0, 'a', hit, id, out);
            emit_read(g, c, ct, hap, f + ins - L, 1, 'b', hit2, id, out);
            id++;
        }
    }
}

Cfg config_for(int id, double scale, Xoshiro256ss& g)
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
    case 3:
        c.ncontig = 24;
        for (int i = 0; i < 24; i++) c.lens.push_back(sc(std::exp(std::log(1e6) + g.uniform() * std::log(20.0))));
        c.paired = true;
        c.gamma_depth = true;
        break;
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
        break;
    default:
        c.lens = {sc(100000)};
    }
    return c;
}

}  // namespace

extern "C" {

// Writes <prefix>.fa, <prefix>.soap and (config 2) <prefix>.snp.  Returns the
// number of alignment lines written, or a negative value on error.  The read
// length of the configuration is returned through *read_len.
long long gsnp_synth_write(int config, unsigned long long seed, double scale, const char* prefix, int* read_len)
{
    Xoshiro256ss g(seed * 0x100000001b3ULL + (uint64_t)config);
    Cfg c = config_for(config, scale, g);
    if (read_len) *read_len = c.read_len;
    std::string pfx(prefix);
    FILE* fa = std::fopen((pfx + ".fa").c_str(), "w");
    FILE* so = std::fopen((pfx + ".soap").c_str(), "w");
    FILE* sn = c.dbsnp ? std::fopen((pfx + ".snp").c_str(), "w") : nullptr;
    if (!fa || !so || (c.dbsnp && !sn)) return -1;
    long long nlines = 0;
    uint32_t id = 0;
    std::vector<char> line(4096);
    for (size_t ci = 0; ci < c.lens.size(); ci++) {
        Contig ct;
        ct.name = c.lens.size() == 1 ? std::string("chr1") : ("chr" + std::to_string(ci + 1));
        make_contig(g, ct, c.lens[ci]);
        if (c.stress) {  // a few N runs in the reference (code 7 marks 'N' when written)
            for (int k = 0; k < 4; k++) {
                uint64_t at = g.below(ct.ref.size() - 2000), run = 50 + g.below(1500);
                for (uint64_t i = at; i < at + run; i++) ct.ref[i] = 7;
            }
        }
        std::fprintf(fa, ">%s\n", ct.name.c_str());
        for (uint64_t i = 0; i < ct.ref.size(); i += 60) {
            uint64_t m = std::min<uint64_t>(60, ct.ref.size() - i);
            for (uint64_t k = 0; k < m; k++) line[k] = ct.ref[i + k] == 7 ? 'N' : kBase[ct.ref[i + k]];
            line[m] = '\n';
            std::fwrite(line.data(), 1, m + 1, fa);
        }
        std::vector<Read> reads;
        if (c.stress) {
            for (uint64_t lo = 0; lo < ct.ref.size(); lo += 5000) {
                uint64_t hi = std::min<uint64_t>(ct.ref.size(), lo + 5000);
                double u = g.uniform();
                double d = u < 0.05 ? 0 : u < 0.12 ? 260 + (double)g.below(40) : u < 0.2 ? 80 + (double)g.below(60) : 5 + 30 * g.uniform();
                reads_for_span(g, c, ct, lo, hi, d, u >= 0.05 && u < 0.09, id, reads);
            }
        } else if (c.gamma_depth) {
            for (uint64_t lo = 0; lo < ct.ref.size(); lo += 10000) {
                uint64_t hi = std::min<uint64_t>(ct.ref.size(), lo + 10000);
                double u = g.uniform();
                if (u < 0.02) continue;  // uncovered block
                if (u < 0.03)
                    reads_for_span(g, c, ct, lo, hi, 200 + (double)g.below(56), true, id, reads);
                else
                    reads_for_span(g, c, ct, lo, hi, g.gamma_int(4, 7.5), false, id, reads);
            }
        } else {
            reads_for_span(g, c, ct, 0, ct.ref.size(), c.depth, false, id, reads);
        }
        std::stable_sort(reads.begin(), reads.end(), [](const Read& a, const Read& b) { return a.pos < b.pos; });
        for (Read& r : reads) {
            if (c.stress && g.uniform() < 0.01) {
                r.indel = g.below(2) ? 100 + 1 + (int)g.below(3) : 200 + 1 + (int)g.below(3);
                r.indel_pos = 5 + (int)g.below(c.read_len - 10);
            }
        }
        for (const Read& r : reads) {
            int mm = 0;
            for (int k = 0; k < c.read_len; k++) {
                char rb = ct.ref[r.pos + k] == 7 ? 'N' : kBase[ct.ref[r.pos + k]];
                mm += r.seq[k] != rb;
            }
            if (mm > 99) mm = 99;
            int n;
            if (r.indel)
                n = std::snprintf(line.data(), line.size(), "r%u/%c\t%s\t%s\t%d\t%c\t%d\t%c\t%s\t%d\t%d\t%d\n", r.id,
                                  r.ab == 'a' ? '1' : '2', r.seq.c_str(), r.qual.c_str(), (int)r.hit, r.ab, c.read_len,
                                  r.strand ? '-' : '+', ct.name.c_str(), r.pos + 1, r.indel, r.indel_pos);
            else
                n = std::snprintf(line.data(), line.size(), "r%u/%c\t%s\t%s\t%d\t%c\t%d\t%c\t%s\t%d\t%d\n", r.id,
                                  r.ab == 'a' ? '1' : '2', r.seq.c_str(), r.qual.c_str(), (int)r.hit, r.ab, c.read_len,
                                  r.strand ? '-' : '+', ct.name.c_str(), r.pos + 1, mm);
            std::fwrite(line.data(), 1, n, so);
            nlines++;
        }
        if (sn) {
            // 1% of positions: half of the true variant sites plus random others;
            // ref + one alt allele, alt frequency U(0.05, 0.5)
            uint64_t want = ct.ref.size() / 100;
            std::vector<uint64_t> pos;
            for (uint64_t i = 0; i < ct.ref.size() && pos.size() < want / 2; i++)
                if ((ct.hap[0][i] != ct.ref[i] || ct.hap[1][i] != ct.ref[i]) && g.uniform() < 0.5) pos.push_back(i);
            while (pos.size() < want) pos.push_back(g.below(ct.ref.size()));
            std::sort(pos.begin(), pos.end());
            pos.erase(std::unique(pos.begin(), pos.end()), pos.end());
            for (uint64_t p : pos) {
                uint8_t r0 = ct.ref[p];
                if (r0 > 3) continue;  // no known SNPs on reference N
                uint8_t alt = ct.hap[0][p] != r0 ? ct.hap[0][p] : (ct.hap[1][p] != r0 ? ct.hap[1][p] : mutate(g, r0));
                double f[4] = {0, 0, 0, 0};
                double af = 0.05 + 0.45 * g.uniform();
                f[alt] = af;
                f[r0] = 1.0 - af;
                int hapmap = g.uniform() < 0.1;
                int valid = g.uniform() < 0.5;
                int indel = g.uniform() < 0.02;
                std::fprintf(sn, "%s\t%llu\t%d\t%d\t%d\t%.6f\t%.6f\t%.6f\t%.6f\trs%llu\n", ct.name.c_str(),
                             (unsigned long long)(p + 1), hapmap, valid, indel, f[0], f[1], f[2], f[3],
                             (unsigned long long)p);
            }
        }
    }
    std::fclose(fa);
    std::fclose(so);
    if (sn) std::fclose(sn);
    return nlines;
}

}  // extern "C"
