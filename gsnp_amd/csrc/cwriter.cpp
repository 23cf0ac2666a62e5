// cwriter.cpp -- GSNP's -C compressed column format (host side).
#include "cwriter.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace gsnp {

namespace {

const char kACTGN[9] = "ACTGNNNN";
const char kAbbv[18] = "AMWRMCYSWYTKRSKGN";

void put(FILE* f, const void* p, size_t sz, size_t n)
{
    if (n) std::fwrite(p, sz, n, f);
}

bool get(FILE* f, void* p, size_t sz, size_t n) { return n == 0 || std::fread(p, sz, n, f) == n; }

// 2-bit words, A0 C1 G2 T3, 16 per word, MSB first, the last word holding
// the remainder right-aligned (kernel/compress_kernel.cu:104-153)
void decode2(const uint32_t* w, uint64_t nw, uint64_t n, char* out)
{
    static const char T[4] = {'A', 'C', 'G', 'T'};
    if (!nw) return;
    for (uint64_t i = 0; i + 1 < nw; i++)
        for (int k = 0; k < 16; k++) out[i * 16 + k] = T[(w[i] >> (30 - 2 * k)) & 3];
    const uint64_t left = n - (nw - 1) * 16;
    for (uint64_t k = 0; k < left; k++) out[(nw - 1) * 16 + k] = T[(w[nw - 1] >> (2 * (left - k - 1))) & 3];
}

template <class T>
void dict_decode(const uint32_t* w, uint64_t nw, const T* dict, uint64_t nd, uint64_t n, T* out)
{
    if (!n || !nw) return;
    const int bits = dict_bits(nd);
    const int per = (int)std::floor(32 / (double)bits);
    const uint32_t mask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1);
    for (uint64_t i = 0; i < nw; i++) {
        const uint64_t m = (i == nw - 1) ? n - (uint64_t)per * i : (uint64_t)per;
        for (uint64_t k = 0; k < m; k++) out[i * per + k] = dict[(w[i] >> (bits * (m - k - 1))) & mask];
    }
}

template <class T>
bool read_rle_column(FILE* f, uint32_t C, T* out)
{
    uint64_t pa[6];
    if (!get(f, pa, 8, 6)) return false;
    std::vector<uint32_t> wv(pa[1]), wl(pa[2]), dl(pa[4]);
    std::vector<T> dv(pa[3]);
    if (!get(f, wv.data(), 4, pa[1]) || !get(f, wl.data(), 4, pa[2]) || !get(f, dv.data(), sizeof(T), pa[3]) ||
        !get(f, dl.data(), 4, pa[4]))
        return false;
    std::vector<T> rv(pa[0]);
    std::vector<uint32_t> rl(pa[0]);
    dict_decode(wv.data(), pa[1], dv.data(), pa[3], pa[0], rv.data());
    dict_decode(wl.data(), pa[2], dl.data(), pa[4], pa[0], rl.data());
    uint64_t k = 0;
    for (uint64_t i = 0; i < pa[0]; i++)
        for (uint32_t m = 0; m < rl[i] && k < C; m++) out[k++] = rv[i];
    return k == C;
}

}  // namespace

int dict_bits(uint64_t n)
{
    int b = (int)std::ceil(std::log((double)n) / std::log((double)2.f));
    return b == 0 ? 1 : b;
}

int CWriter::open(const std::string& prefix, std::string& err)
{
    prefix_ = prefix;
    f346_ = std::fopen((prefix + ".3_4_6").c_str(), "wb");
    f5_ = std::fopen((prefix + ".5_7_8_9_14_16").c_str(), "wb");
    f10_ = std::fopen((prefix + ".10").c_str(), "wb");
    f11_ = std::fopen((prefix + ".11_12_13").c_str(), "wb");
    f15_ = std::fopen((prefix + ".15_17").c_str(), "wb");
    if (!f346_ || !f5_ || !f10_ || !f11_ || !f15_) {
        err = "cannot create the -C files " + prefix + ".*";
        return 1;
    }
    return 0;
}

// The chunk of an empty window (a chromosome whose length is a multiple of
// the window ends with one, cpp/call_genotype.cc:636-644): an all-N chunk of
// size 0 in every file (cpp/compress.h:151-1312).
void CWriter::write_empty_window(const std::string& name, uint32_t ws)
{
    names_.push_back(name);
    starts_.push_back((uint64_t)ws + 1);
    sizes_.push_back(0);
    const uint64_t param[7] = {1, 0, 0, 0, 0, 0, 0};
    put(f346_, param, 8, 7);
    const int32_t nind = 1;
    for (FILE* f : {f5_, f10_, f11_, f15_}) put(f, &nind, 4, 1);
}

void CWriter::write_chunks(const std::string& name, const ColumnPass& ch)
{
    for (uint32_t i = 0; i < ch.n_windows; i++) {
        names_.push_back(name);
        starts_.push_back(ch.starts[i]);
        sizes_.push_back(ch.sizes[i]);
    }
    FILE* fs[5] = {f346_, f5_, f10_, f11_, f15_};
    for (int f = 0; f < 5; f++) put(fs[f], ch.file[f], 1, ch.bytes[f]);
}

int CWriter::close(std::string& err)
{
    for (FILE* f : {f346_, f5_, f10_, f11_, f15_})
        if (f) std::fclose(f);
    f346_ = f5_ = f10_ = f11_ = f15_ = nullptr;
    FILE* f = std::fopen((prefix_ + ".1_2").c_str(), "w");
    if (!f) {
        err = "cannot create " + prefix_ + ".1_2";
        return 1;
    }
    std::fprintf(f, "%lu\n", (unsigned long)names_.size());
    for (size_t i = 0; i < names_.size(); i++)
        std::fprintf(f, "%s\t%lu\t%lu\n", names_[i].c_str(), (unsigned long)starts_[i], (unsigned long)sizes_[i]);
    std::fclose(f);
    return 0;
}

ColumnReader::~ColumnReader()
{
    for (FILE* f : {f12_, f346_, f5_, f10_, f11_, f15_})
        if (f) std::fclose(f);
}

int ColumnReader::open(const std::string& prefix, std::string& err)
{
    f12_ = std::fopen((prefix + ".1_2").c_str(), "r");
    if (!f12_) {
        err = "cannot open " + prefix + ".1_2";
        return 1;
    }
    if (std::fscanf(f12_, "%lu", &nwin_) != 1) {
        err = "bad window index";
        return 2;
    }
    f346_ = std::fopen((prefix + ".3_4_6").c_str(), "rb");
    f5_ = std::fopen((prefix + ".5_7_8_9_14_16").c_str(), "rb");
    f10_ = std::fopen((prefix + ".10").c_str(), "rb");
    f11_ = std::fopen((prefix + ".11_12_13").c_str(), "rb");
    f15_ = std::fopen((prefix + ".15_17").c_str(), "rb");
    if (!f346_ || !f5_ || !f10_ || !f11_ || !f15_) {
        err = "cannot open the -C files";
        return 1;
    }
    return 0;
}

// One window's chunks back to its 17 columns (the read_decompress calls of
// decompressAll, cpp/decompression.cpp:130-255).
int ColumnReader::next(CnsWindow& w, bool& got, std::string& err)
{
    got = false;
    if (wi_ >= nwin_) return 0;
    char name[1024];
    unsigned long ps, C;
    if (std::fscanf(f12_, "%1023s\t%lu\t%lu\n", name, &ps, &C) != 3) {
        err = "bad window index row";
        return 2;
    }
    wi_++;
    w.chr = name;
    w.start = ps;
    w.n = C;
    w.c3.assign(C, 'N'), w.c4.assign(C, 'N'), w.c6.assign(C, 'N'), w.c10.assign(C, 'N');
    w.c5.assign(C, 0), w.c7.assign(C, 0), w.c8.assign(C, 0), w.c9.assign(C, 0), w.c14.assign(C, 0);
    w.c11.assign(C, 0), w.c12.assign(C, 0), w.c13.assign(C, 0), w.c17.assign(C, 0);
    w.c15.assign(C, 1.0), w.c16.assign(C, 255.0);
    uint64_t pa[7];
    bool ok = get(f346_, pa, 8, 7);
    if (ok && pa[0] != 1) {
        std::vector<int32_t> id4(pa[5]), id6(pa[6]), idn(pa[4]);
        std::vector<char> v4(pa[5]), v6(pa[6]);
        std::vector<uint32_t> code(pa[3]);
        ok = get(f346_, id4.data(), 4, pa[5]) && get(f346_, v4.data(), 1, pa[5]) && get(f346_, id6.data(), 4, pa[6]) &&
             get(f346_, v6.data(), 1, pa[6]) && get(f346_, idn.data(), 4, pa[4]) && get(f346_, code.data(), 4, pa[3]);
        if (ok) {
            decode2(code.data(), pa[3], C, w.c3.data());
            for (int32_t i : idn) w.c3[i] = 'N';
            w.c4 = w.c3;
            for (size_t i = 0; i < id4.size(); i++) w.c4[id4[i]] = v4[i];
            w.c6 = w.c3;
            for (size_t i = 0; i < id6.size(); i++) w.c6[id6[i]] = v6[i];
        }
    }
    int32_t nind = 1;
    ok = ok && get(f5_, &nind, 4, 1);
    if (ok && nind != 1)
        ok = read_rle_column(f5_, C, w.c5.data()) && read_rle_column(f5_, C, w.c7.data()) &&
             read_rle_column(f5_, C, w.c8.data()) && read_rle_column(f5_, C, w.c9.data()) &&
             read_rle_column(f5_, C, w.c14.data()) && read_rle_column(f5_, C, w.c16.data());
    ok = ok && get(f10_, &nind, 4, 1);
    if (ok && nind != 1) {
        uint64_t p3[3];
        ok = get(f10_, p3, 8, 3);
        std::vector<uint32_t> code(ok ? p3[1] : 0), ni(ok ? p3[0] : 0);
        ok = ok && get(f10_, code.data(), 4, p3[1]) && get(f10_, ni.data(), 4, p3[0]);
        if (ok) {
            decode2(code.data(), p3[1], C, w.c10.data());
            for (uint32_t i : ni) w.c10[i] = 'N';
        }
    }
    ok = ok && get(f11_, &nind, 4, 1);
    if (ok && nind != 1) {
        uint64_t p4[4];
        ok = get(f11_, p4, 8, 4);
        const uint64_t nc = ok ? p4[1] : 0, n12 = ok ? p4[2] : 0, n13 = ok ? p4[3] : 0;
        std::vector<int32_t> x12i(n12), x12v(n12), x13i(n13), x13v(n13), cid(nc);
        ok = ok && get(f11_, x12i.data(), 4, n12) && get(f11_, x12v.data(), 4, n12) &&
             get(f11_, x13i.data(), 4, n13) && get(f11_, x13v.data(), 4, n13) && get(f11_, cid.data(), 4, nc);
        std::vector<int32_t> vals[3];
        for (int k = 0; k < 3 && ok; k++) {
            uint64_t dp[2];
            ok = get(f11_, dp, 8, 2);
            std::vector<uint32_t> words(ok ? dp[0] : 0);
            std::vector<int32_t> dict(ok ? dp[1] : 0);
            ok = ok && get(f11_, words.data(), 4, dp[0]) && get(f11_, dict.data(), 4, dp[1]);
            vals[k].assign(nc, 0);
            if (ok) dict_decode(words.data(), dp[0], dict.data(), dp[1], nc, vals[k].data());
        }
        if (ok) {
            for (uint64_t i = 0; i < nc; i++)
                w.c11[cid[i]] = vals[0][i], w.c12[cid[i]] = vals[1][i], w.c13[cid[i]] = vals[2][i];
            for (uint64_t i = 0; i < n12; i++) w.c12[x12i[i]] = x12v[i];
            for (uint64_t i = 0; i < n13; i++) w.c13[x13i[i]] = x13v[i];
        }
    }
    ok = ok && get(f15_, &nind, 4, 1);
    if (ok && nind != 1) {
        int32_t n = 0;
        ok = get(f15_, &n, 4, 1);
        std::vector<int32_t> id(ok ? n : 0);
        std::vector<double> val(ok ? n : 0);
        ok = ok && get(f15_, id.data(), 4, n) && get(f15_, val.data(), 8, n);
        if (ok)
            for (int32_t i = 0; i < n; i++) w.c15[id[i]] = val[i];
        ok = ok && get(f15_, &n, 4, 1);
        if (ok) {
            id.assign(n, 0);
            ok = get(f15_, id.data(), 4, n);
            if (ok)
                for (int32_t i = 0; i < n; i++) w.c17[id[i]] = 1;
        }
    }
    if (!ok) {
        err = "truncated -C column file";
        return 2;
    }
    got = true;
    return 0;
}

int decompress_files(const std::string& prefix, const std::string& out, std::string& err)
{
    ColumnReader rd;
    int rc = rd.open(prefix, err);
    if (rc) return rc;
    FILE* fo = std::fopen(out.c_str(), "w");
    if (!fo) {
        err = "cannot create " + out;
        return 1;
    }
    CnsWindow w;
    bool got = true;
    while (!rc) {
        rc = rd.next(w, got, err);
        if (rc || !got) break;
        const char* name = w.chr.c_str();
        for (uint64_t i = 0; i < w.n; i++) {
            const unsigned long pos = (unsigned long)(w.start + i);
            if (w.c3[i] == 'N' && w.c4[i] == 'N' && w.c6[i] == 'N')
                std::fprintf(fo, "%s\t%lu\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n", name, pos);
            else
                std::fprintf(fo, "%s\t%lu\t%c\t%c\t%d\t%c\t%d\t%d\t%d\t%c\t%d\t%d\t%d\t%d\t%#g\t%#g\t%d\n", name, pos,
                             w.c3[i], w.c4[i], w.c5[i], w.c6[i], w.c7[i], w.c8[i], w.c9[i], w.c10[i], w.c11[i], w.c12[i],
                             w.c13[i], w.c14[i], w.c15[i], w.c16[i], w.c17[i]);
        }
    }
    std::fclose(fo);
    return rc;
}

}  // namespace gsnp
