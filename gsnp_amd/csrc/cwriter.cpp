// cwriter.cpp -- GSNP's -C compressed column format (host side).
#include "cwriter.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/gsnp_abi.h"

namespace gsnp {

namespace {

const char kACTGN[9] = "ACTGNNNN";
const char kAbbv[18] = "AMWRMCYSWYTKRSKGN";

bool put(FILE* f, const void* p, size_t sz, size_t n) { return n == 0 || std::fwrite(p, sz, n, f) == n; }

bool get(FILE* f, void* p, size_t sz, size_t n) { return n == 0 || std::fread(p, sz, n, f) == n; }

// 2-bit words, A0 C1 G2 T3, 16 per word, MSB first, the last word holding
// the remainder right-aligned (kernel/compress_kernel.cu:104-153); nw must be
// ceil(n / 16)
void decode2(const uint32_t* w, uint64_t nw, uint64_t n, char* out)
{
    static const char T[4] = {'A', 'C', 'G', 'T'};
    if (!nw) return;
    for (uint64_t i = 0; i + 1 < nw; i++)
        for (int k = 0; k < 16; k++) out[i * 16 + k] = T[(w[i] >> (30 - 2 * k)) & 3];
    const uint64_t left = n - (nw - 1) * 16;
    for (uint64_t k = 0; k < left; k++) out[(nw - 1) * 16 + k] = T[(w[nw - 1] >> (2 * (left - k - 1))) & 3];
}

// Decodes n values; false when the word count does not fit n and nd or an
// index points past the dictionary.
template <class T>
bool dict_decode(const uint32_t* w, uint64_t nw, const T* dict, uint64_t nd, uint64_t n, T* out)
{
    if (!n) return nw == 0;
    if (nd < 1 || nd > n || nw != dict_words(n, nd)) return false;
    const int bits = dict_bits(nd);
    const int per = (int)std::floor(32 / (double)bits);
    const uint32_t mask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1);
    for (uint64_t i = 0; i < nw; i++) {
        const uint64_t m = (i == nw - 1) ? n - (uint64_t)per * i : (uint64_t)per;
        for (uint64_t k = 0; k < m; k++) {
            const uint32_t idx = (w[i] >> (bits * (m - k - 1))) & mask;
            if (idx >= nd) return false;
            out[i * per + k] = dict[idx];
        }
    }
    return true;
}

// One run-length + dictionary column of C values (Column_RLEDict,
// cpp/compress.h:501-783), read and checked: every count against C before
// anything is allocated, every dictionary index, and runs covering exactly C
// values.
bool read_rle_column(FILE* f, uint64_t C, bool f64, RawRle& r, std::string& why)
{
    uint64_t pa[6];
    if (!get(f, pa, 8, 6)) return why = "truncated run-length column", false;
    const uint64_t nr = pa[0], vs = f64 ? 8 : 4;
    if (nr < 1 || nr > C || pa[3] < 1 || pa[3] > nr || pa[4] < 1 || pa[4] > nr || pa[1] != dict_words(nr, pa[3]) ||
        pa[2] != dict_words(nr, pa[4]) || pa[5] != 4 * pa[1] + 4 * pa[2] + vs * pa[3] + 4 * pa[4])
        return why = "inconsistent run-length column header", false;
    r.nrun = nr;
    r.nd_val = pa[3];
    r.nd_len = pa[4];
    r.wv.resize(pa[1]);
    r.wl.resize(pa[2]);
    r.dl.resize(pa[4]);
    r.dv32.resize(f64 ? 0 : pa[3]);
    r.dv64.resize(f64 ? pa[3] : 0);
    void* dv = f64 ? (void*)r.dv64.data() : (void*)r.dv32.data();
    if (!get(f, r.wv.data(), 4, pa[1]) || !get(f, r.wl.data(), 4, pa[2]) || !get(f, dv, vs, pa[3]) ||
        !get(f, r.dl.data(), 4, pa[4]))
        return why = "truncated run-length column", false;
    // indices inside the dictionaries, lengths adding up to C
    std::vector<uint32_t> rl(nr);
    std::vector<int32_t> scratch32(f64 ? 0 : nr);
    std::vector<double> scratch64(f64 ? nr : 0);
    const bool ok = (f64 ? dict_decode(r.wv.data(), pa[1], r.dv64.data(), pa[3], nr, scratch64.data())
                         : dict_decode(r.wv.data(), pa[1], r.dv32.data(), pa[3], nr, scratch32.data())) &&
                    dict_decode(r.wl.data(), pa[2], r.dl.data(), pa[4], nr, rl.data());
    if (!ok) return why = "bad dictionary index in a run-length column", false;
    uint64_t k = 0;
    for (uint64_t i = 0; i < nr; i++) {
        if (rl[i] > C - k) return why = "runs longer than the window", false;
        k += rl[i];
    }
    if (k != C) return why = "runs shorter than the window", false;
    return true;
}

// The C values of a checked run-length column.
template <class T>
void expand_rle(const RawRle& r, const T* dv, T* out)
{
    std::vector<T> rv(r.nrun);
    std::vector<uint32_t> rl(r.nrun);
    dict_decode(r.wv.data(), r.wv.size(), dv, r.nd_val, r.nrun, rv.data());
    dict_decode(r.wl.data(), r.wl.size(), r.dl.data(), r.nd_len, r.nrun, rl.data());
    uint64_t k = 0;
    for (uint64_t i = 0; i < r.nrun; i++)
        for (uint32_t m = 0; m < rl[i]; m++) out[k++] = rv[i];
}

bool ids_in_window(const std::vector<int32_t>& id, uint64_t C)
{
    for (int32_t i : id)
        if (i < 0 || (uint64_t)i >= C) return false;
    return true;
}

}  // namespace

int dict_bits(uint64_t n)
{
    int b = (int)std::ceil(std::log((double)n) / std::log((double)2.f));
    return b == 0 ? 1 : b;
}

// Words that n dictionary indices over nd entries take (cpp/dict_compression.h:116-166);
// n == 0 takes none.  nd must be in [1, n] when n > 0.
uint64_t dict_words(uint64_t n, uint64_t nd)
{
    if (!n) return 0;
    const int per = (int)std::floor(32 / (double)dict_bits(nd));
    return (n + per - 1) / per;
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
int CWriter::write_empty_window(const std::string& name, uint32_t ws, std::string& err)
{
    names_.push_back(name);
    starts_.push_back((uint64_t)ws + 1);
    sizes_.push_back(0);
    const uint64_t param[7] = {1, 0, 0, 0, 0, 0, 0};
    bool ok = put(f346_, param, 8, 7);
    const int32_t nind = 1;
    for (FILE* f : {f5_, f10_, f11_, f15_}) ok = put(f, &nind, 4, 1) && ok;
    if (!ok) err = "cannot write the -C files " + prefix_ + ".*";
    return ok ? 0 : 1;
}

int CWriter::write_chunks(const std::string& name, const ColumnPass& ch, std::string& err)
{
    for (uint32_t i = 0; i < ch.n_windows; i++) {
        names_.push_back(name);
        starts_.push_back(ch.starts[i]);
        sizes_.push_back(ch.sizes[i]);
    }
    FILE* fs[5] = {f346_, f5_, f10_, f11_, f15_};
    bool ok = true;
    for (int f = 0; f < 5; f++) ok = put(fs[f], ch.file[f], 1, ch.bytes[f]) && ok;
    if (!ok) err = "cannot write the -C files " + prefix_ + ".*";
    return ok ? 0 : 1;
}

int CWriter::close(std::string& err)
{
    bool ok = true;
    for (FILE* f : {f346_, f5_, f10_, f11_, f15_})
        if (f) ok = std::fclose(f) == 0 && ok;
    f346_ = f5_ = f10_ = f11_ = f15_ = nullptr;
    if (!ok) {
        err = "cannot write the -C files " + prefix_ + ".*";
        return 1;
    }
    FILE* f = std::fopen((prefix_ + ".1_2").c_str(), "w");
    if (!f) {
        err = "cannot create " + prefix_ + ".1_2";
        return 1;
    }
    ok = std::fprintf(f, "%lu\n", (unsigned long)names_.size()) > 0;
    for (size_t i = 0; i < names_.size() && ok; i++)
        ok = std::fprintf(f, "%s\t%lu\t%lu\n", names_[i].c_str(), (unsigned long)starts_[i], (unsigned long)sizes_[i]) > 0;
    ok = std::fclose(f) == 0 && ok;
    if (!ok) {
        err = "cannot write " + prefix_ + ".1_2";
        return 1;
    }
    return 0;
}

CWriter::~CWriter()
{
    for (FILE* f : {f346_, f5_, f10_, f11_, f15_})
        if (f) std::fclose(f);
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

// One window's chunks, read and checked (the reads of the read_decompress
// calls of decompressAll, cpp/decompression.cpp:130-255).  Every count and
// index read from the files is checked against the window size before it is
// used, so a corrupt or truncated file gives an error, never an out-of-range
// access.
int ColumnReader::read_chunk(RawChunk& c, bool& got, std::string& err)
{
    got = false;
    if (wi_ >= nwin_) return 0;
    char name[1024];
    unsigned long ps, C;
    if (std::fscanf(f12_, "%1023s\t%lu\t%lu\n", name, &ps, &C) != 3) {
        err = "bad window index row";
        return 2;
    }
    if (C > (unsigned long)GSNP_WIN) {
        err = "window size in the index above 262144";
        return 2;
    }
    wi_++;
    c = RawChunk();
    c.chr = name;
    c.start = ps;
    c.n = C;
    std::string why;
    auto bad = [&](const char* file, const std::string& what) {
        err = std::string(file) + ": " + what + " (window " + std::to_string(wi_ - 1) + ")";
        return 2;
    };
    // .3_4_6 (cpp/compress.h:187-409): {isAllN, dataSize, winSize, codeLen, nN, n4, n6}
    uint64_t pa[7];
    if (!get(f346_, pa, 8, 7)) return bad(".3_4_6", "truncated");
    if (pa[0] > 1 || pa[2] != C) return bad(".3_4_6", "header does not match the window index");
    c.n346 = pa[0] == 1;
    if (!c.n346) {
        if (pa[3] != (C + 15) / 16 || pa[4] > C || pa[5] > C || pa[6] > C ||
            pa[1] != 5 * pa[5] + 5 * pa[6] + 4 * pa[4] + 4 * pa[3])
            return bad(".3_4_6", "inconsistent header");
        c.id4.resize(pa[5]), c.v4.resize(pa[5]), c.id6.resize(pa[6]), c.v6.resize(pa[6]), c.idn.resize(pa[4]);
        c.code3.resize(pa[3]);
        if (!get(f346_, c.id4.data(), 4, pa[5]) || !get(f346_, c.v4.data(), 1, pa[5]) ||
            !get(f346_, c.id6.data(), 4, pa[6]) || !get(f346_, c.v6.data(), 1, pa[6]) ||
            !get(f346_, c.idn.data(), 4, pa[4]) || !get(f346_, c.code3.data(), 4, pa[3]))
            return bad(".3_4_6", "truncated");
        if (!ids_in_window(c.id4, C) || !ids_in_window(c.id6, C) || !ids_in_window(c.idn, C))
            return bad(".3_4_6", "site index outside the window");
    } else if (pa[1] != 0) {
        return bad(".3_4_6", "all-N chunk with data");
    }
    // .5_7_8_9_14_16 (cpp/compress.h:501-783)
    int32_t nind = 1;
    if (!get(f5_, &nind, 4, 1)) return bad(".5_7_8_9_14_16", "truncated");
    if (nind != 0 && nind != 1) return bad(".5_7_8_9_14_16", "bad N indicator");
    c.n5 = nind == 1;
    for (int k = 0; k < 6 && !c.n5; k++)
        if (!read_rle_column(f5_, C, k == 5, c.rle[k], why)) return bad(".5_7_8_9_14_16", why);
    // .10 (cpp/compress.h:786-897): {numN, numWord, numElement}, words, N indices
    if (!get(f10_, &nind, 4, 1)) return bad(".10", "truncated");
    if (nind != 0 && nind != 1) return bad(".10", "bad N indicator");
    c.n10 = nind == 1;
    if (!c.n10) {
        uint64_t p3[3];
        if (!get(f10_, p3, 8, 3)) return bad(".10", "truncated");
        if (p3[2] != C || p3[1] != (C + 15) / 16 || p3[0] > C) return bad(".10", "inconsistent header");
        c.code10.resize(p3[1]);
        c.ni10.resize(p3[0]);
        if (!get(f10_, c.code10.data(), 4, p3[1]) || !get(f10_, c.ni10.data(), 4, p3[0])) return bad(".10", "truncated");
        for (uint32_t i : c.ni10)
            if (i >= C) return bad(".10", "site index outside the window");
    }
    // .11_12_13 (cpp/compress.h:951-1175): {winSize, numCommon, numExtra12,
    // numExtra13}, extra lists, common ids, three dictionary-coded columns
    if (!get(f11_, &nind, 4, 1)) return bad(".11_12_13", "truncated");
    if (nind != 0 && nind != 1) return bad(".11_12_13", "bad N indicator");
    c.n11 = nind == 1;
    if (!c.n11) {
        uint64_t p4[4];
        if (!get(f11_, p4, 8, 4)) return bad(".11_12_13", "truncated");
        if (p4[0] != C || p4[1] > C || p4[2] > C || p4[3] > C) return bad(".11_12_13", "inconsistent header");
        const uint64_t nc = p4[1], n12 = p4[2], n13 = p4[3];
        c.x12i.resize(n12), c.x12v.resize(n12), c.x13i.resize(n13), c.x13v.resize(n13), c.cid.resize(nc);
        if (!get(f11_, c.x12i.data(), 4, n12) || !get(f11_, c.x12v.data(), 4, n12) ||
            !get(f11_, c.x13i.data(), 4, n13) || !get(f11_, c.x13v.data(), 4, n13) || !get(f11_, c.cid.data(), 4, nc))
            return bad(".11_12_13", "truncated");
        if (!ids_in_window(c.x12i, C) || !ids_in_window(c.x13i, C) || !ids_in_window(c.cid, C))
            return bad(".11_12_13", "site index outside the window");
        for (int k = 0; k < 3; k++) {
            uint64_t dp[2];
            if (!get(f11_, dp, 8, 2)) return bad(".11_12_13", "truncated");
            if (dp[1] > nc || (nc && dp[1] < 1) || dp[0] != dict_words(nc, dp[1]))
                return bad(".11_12_13", "inconsistent dictionary header");
            c.w11[k].resize(dp[0]);
            c.d11[k].resize(dp[1]);
            c.nd11[k] = dp[1];
            if (!get(f11_, c.w11[k].data(), 4, dp[0]) || !get(f11_, c.d11[k].data(), 4, dp[1]))
                return bad(".11_12_13", "truncated");
            std::vector<int32_t> vals(nc);
            if (!dict_decode(c.w11[k].data(), dp[0], c.d11[k].data(), dp[1], nc, vals.data()))
                return bad(".11_12_13", "bad dictionary index");
        }
    }
    // .15_17 (cpp/compress.h:1178-1312): rank sums != 1, then known-SNP flags
    if (!get(f15_, &nind, 4, 1)) return bad(".15_17", "truncated");
    if (nind != 0 && nind != 1) return bad(".15_17", "bad N indicator");
    c.n15 = nind == 1;
    if (!c.n15) {
        int32_t n = 0;
        if (!get(f15_, &n, 4, 1)) return bad(".15_17", "truncated");
        if (n < 0 || (uint64_t)n > C) return bad(".15_17", "inconsistent count");
        c.id15.resize(n);
        c.v15.resize(n);
        if (!get(f15_, c.id15.data(), 4, n) || !get(f15_, c.v15.data(), 8, n)) return bad(".15_17", "truncated");
        if (!ids_in_window(c.id15, C)) return bad(".15_17", "site index outside the window");
        if (!get(f15_, &n, 4, 1)) return bad(".15_17", "truncated");
        if (n < 0 || (uint64_t)n > C) return bad(".15_17", "inconsistent count");
        c.id17.resize(n);
        if (!get(f15_, c.id17.data(), 4, n)) return bad(".15_17", "truncated");
        if (!ids_in_window(c.id17, C)) return bad(".15_17", "site index outside the window");
    }
    got = true;
    return 0;
}

int decode_chunk(const RawChunk& c, CnsWindow& w, std::string& err)
{
    (void)err;
    const uint64_t C = c.n;
    w.chr = c.chr;
    w.start = c.start;
    w.n = C;
    w.c3.assign(C, 'N'), w.c4.assign(C, 'N'), w.c6.assign(C, 'N'), w.c10.assign(C, 'N');
    w.c5.assign(C, 0), w.c7.assign(C, 0), w.c8.assign(C, 0), w.c9.assign(C, 0), w.c14.assign(C, 0);
    w.c11.assign(C, 0), w.c12.assign(C, 0), w.c13.assign(C, 0), w.c17.assign(C, 0);
    w.c15.assign(C, 1.0), w.c16.assign(C, 255.0);
    if (!c.n346) {
        decode2(c.code3.data(), c.code3.size(), C, w.c3.data());
        for (int32_t i : c.idn) w.c3[i] = 'N';
        w.c4 = w.c3;
        for (size_t i = 0; i < c.id4.size(); i++) w.c4[c.id4[i]] = c.v4[i];
        w.c6 = w.c3;
        for (size_t i = 0; i < c.id6.size(); i++) w.c6[c.id6[i]] = c.v6[i];
    }
    if (!c.n5) {
        int32_t* cols[5] = {w.c5.data(), w.c7.data(), w.c8.data(), w.c9.data(), w.c14.data()};
        for (int k = 0; k < 5; k++) expand_rle(c.rle[k], c.rle[k].dv32.data(), cols[k]);
        expand_rle(c.rle[5], c.rle[5].dv64.data(), w.c16.data());
    }
    if (!c.n10) {
        decode2(c.code10.data(), c.code10.size(), C, w.c10.data());
        for (uint32_t i : c.ni10) w.c10[i] = 'N';
    }
    if (!c.n11) {
        const uint64_t nc = c.cid.size();
        std::vector<int32_t> vals[3];
        for (int k = 0; k < 3; k++) {
            vals[k].assign(nc, 0);
            dict_decode(c.w11[k].data(), c.w11[k].size(), c.d11[k].data(), c.nd11[k], nc, vals[k].data());
        }
        for (uint64_t i = 0; i < nc; i++)
            w.c11[c.cid[i]] = vals[0][i], w.c12[c.cid[i]] = vals[1][i], w.c13[c.cid[i]] = vals[2][i];
        for (size_t i = 0; i < c.x12i.size(); i++) w.c12[c.x12i[i]] = c.x12v[i];
        for (size_t i = 0; i < c.x13i.size(); i++) w.c13[c.x13i[i]] = c.x13v[i];
    }
    if (!c.n15) {
        for (size_t i = 0; i < c.id15.size(); i++) w.c15[c.id15[i]] = c.v15[i];
        for (int32_t i : c.id17) w.c17[i] = 1;
    }
    return 0;
}

// One window's chunks back to its 17 columns.
int ColumnReader::next(CnsWindow& w, bool& got, std::string& err)
{
    RawChunk c;
    int rc = read_chunk(c, got, err);
    if (rc || !got) return rc;
    return decode_chunk(c, w, err);
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
    bool got = true, ok = true;
    while (!rc && ok) {
        rc = rd.next(w, got, err);
        if (rc || !got) break;
        const char* name = w.chr.c_str();
        for (uint64_t i = 0; i < w.n && ok; i++) {
            const unsigned long pos = (unsigned long)(w.start + i);
            if (w.c3[i] == 'N' && w.c4[i] == 'N' && w.c6[i] == 'N')
                ok = std::fprintf(fo, "%s\t%lu\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n", name, pos) > 0;
            else
                ok = std::fprintf(fo, "%s\t%lu\t%c\t%c\t%d\t%c\t%d\t%d\t%d\t%c\t%d\t%d\t%d\t%d\t%#g\t%#g\t%d\n", name,
                                  pos, w.c3[i], w.c4[i], w.c5[i], w.c6[i], w.c7[i], w.c8[i], w.c9[i], w.c10[i], w.c11[i],
                                  w.c12[i], w.c13[i], w.c14[i], w.c15[i], w.c16[i], w.c17[i]) > 0;
        }
    }
    ok = std::fclose(fo) == 0 && ok;
    if (!rc && !ok) {
        err = "cannot write " + out;
        rc = 1;
    }
    return rc;
}

}  // namespace gsnp
