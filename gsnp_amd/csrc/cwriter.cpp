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

// 2-bit codes A0 C1 G2 T3, 16 per word, MSB first, last word right-aligned
// (kernel/compress_kernel.cu:85-153); anything else encodes as A.
uint32_t code2(char b)
{
    switch (b) {
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 0;
    }
}

std::vector<uint32_t> encode2(const std::vector<char>& s)
{
    const uint64_t n = s.size();
    const uint64_t nw = (uint64_t)std::ceil(n / 16.0);
    std::vector<uint32_t> w(nw, 0);
    for (uint64_t i = 0; i < nw; i++) {
        const uint64_t m = (i == nw - 1) ? n - i * 16 : 16;
        uint32_t x = 0;
        for (uint64_t k = 0; k < m; k++) x = (x << 2) | code2(s[i * 16 + k]);
        w[i] = x;
    }
    return w;
}

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
void rle(const std::vector<T>& v, std::vector<T>& val, std::vector<uint32_t>& len)
{
    val.clear();
    len.clear();
    for (size_t i = 0; i < v.size(); i++) {
        if (!val.empty() && val.back() == v[i])
            len.back()++;
        else
            val.push_back(v[i]), len.push_back(1);
    }
}

template <class T>
void write_rle_column(FILE* f, const std::vector<T>& col)
{
    std::vector<T> rv;
    std::vector<uint32_t> rl;
    rle(col, rv, rl);
    std::vector<uint32_t> wv, wl;
    std::vector<T> dv;
    std::vector<uint32_t> dl;
    const uint64_t nwv = dict_encode(rv.data(), rv.size(), wv, dv);
    const uint64_t nwl = dict_encode(rl.data(), rl.size(), wl, dl);
    const uint64_t param[6] = {rv.size(), nwv, nwl, dv.size(), dl.size(),
                               4 * nwv + 4 * nwl + sizeof(T) * dv.size() + 4 * dl.size()};
    put(f, param, 8, 6);
    put(f, wv.data(), 4, nwv);
    put(f, wl.data(), 4, nwl);
    put(f, dv.data(), sizeof(T), dv.size());
    put(f, dl.data(), 4, dl.size());
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

template <class T>
uint64_t dict_encode(const T* in, uint64_t n, std::vector<uint32_t>& words, std::vector<T>& dict)
{
    words.clear();
    dict.assign(in, in + n);
    if (!n) return 0;  // empty column: no words, empty dictionary
    std::sort(dict.begin(), dict.end());
    dict.erase(std::unique(dict.begin(), dict.end()), dict.end());
    const int bits = dict_bits(dict.size());
    const uint64_t per = (uint64_t)std::floor(32 / (double)bits);
    const uint64_t nw = (uint64_t)std::ceil(n / (double)per);
    words.assign(nw, 0);
    for (uint64_t i = 0; i < n; i++) {
        const uint32_t idx = (uint32_t)(std::lower_bound(dict.begin(), dict.end(), in[i]) - dict.begin());
        uint32_t& w = words[i / per];
        w = bits >= 32 ? 0u : (w << bits);
        w |= idx;
    }
    return nw;
}

template uint64_t dict_encode<int32_t>(const int32_t*, uint64_t, std::vector<uint32_t>&, std::vector<int32_t>&);
template uint64_t dict_encode<uint32_t>(const uint32_t*, uint64_t, std::vector<uint32_t>&, std::vector<uint32_t>&);
template uint64_t dict_encode<double>(const double*, uint64_t, std::vector<uint32_t>&, std::vector<double>&);

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

void CWriter::write_window(const std::string& name, uint32_t ws, const WindowCols& w)
{
    const uint32_t C = w.n;
    names_.push_back(name);
    starts_.push_back((uint64_t)ws + 1);
    sizes_.push_back(C);
    bool all_n = true;
    for (uint32_t j = 0; j < C && all_n; j++) all_n = w.nflag(j);
    // columns 3, 4, 6 (cpp/compress.h:187-409)
    if (all_n) {
        const uint64_t param[7] = {1, 0, C, 0, 0, 0, 0};
        put(f346_, param, 8, 7);
    } else {
        std::vector<char> c3(C), c4(C), c6(C);
        for (uint32_t j = 0; j < C; j++) {
            if (w.nrow(j)) {
                c3[j] = c4[j] = c6[j] = 'N';
            } else {
                c3[j] = kACTGN[w.ori[j] & 7];
                c4[j] = kAbbv[w.type1[j]];
                c6[j] = kACTGN[w.base1[j]];
            }
        }
        std::vector<int32_t> id4, id6, idn;
        std::vector<char> v4, v6;
        for (uint32_t j = 0; j < C; j++) {
            if (c4[j] != c3[j]) id4.push_back((int32_t)j), v4.push_back(c4[j]);
            if (c6[j] != c3[j]) id6.push_back((int32_t)j), v6.push_back(c6[j]);
            if (c3[j] == 'N') idn.push_back((int32_t)j), c3[j] = 'A';
        }
        std::vector<uint32_t> code = encode2(c3);
        const uint64_t size = 5 * id4.size() + 5 * id6.size() + 4 * idn.size() + 4 * code.size();
        const uint64_t param[7] = {0, size, C, code.size(), idn.size(), id4.size(), id6.size()};
        put(f346_, param, 8, 7);
        put(f346_, id4.data(), 4, id4.size());
        put(f346_, v4.data(), 1, v4.size());
        put(f346_, id6.data(), 4, id6.size());
        put(f346_, v6.data(), 1, v6.size());
        put(f346_, idn.data(), 4, idn.size());
        put(f346_, code.data(), 4, code.size());
    }
    const int32_t nind = all_n ? 1 : 0;
    // columns 5, 7, 8, 9, 14, 16: run lengths + dictionaries (cpp/compress.h:501-610)
    put(f5_, &nind, 4, 1);
    if (!all_n) {
        std::vector<int32_t> col(C);
        for (int which = 0; which < 5; which++) {
            for (uint32_t j = 0; j < C; j++) {
                int32_t v = 0;
                if (!w.nrow(j)) {
                    const uint32_t b = w.base1[j];
                    const int qsb = w.qs[4ull * j + b], cub = w.cu[4ull * j + b];
                    switch (which) {
                    case 0: v = w.q_cns[j]; break;
                    case 1: v = qsb == 0 ? 0 : qsb / cub; break;
                    case 2: v = cub; break;
                    case 3: v = w.ca[4ull * j + b]; break;
                    default: v = w.depth[j]; break;
                    }
                }
                col[j] = v;
            }
            write_rle_column(f5_, col);
        }
        std::vector<double> cn(C);
        for (uint32_t j = 0; j < C; j++)
            cn[j] = w.nrow(j) ? 255.0 : (w.depth[j] == 0 ? 255 : (double)w.repeat[j] / w.depth[j]);
        write_rle_column(f5_, cn);
    }
    // column 10 (cpp/compress.h:786-897)
    put(f10_, &nind, 4, 1);
    if (!all_n) {
        std::vector<char> c10(C);
        std::vector<uint32_t> nidx;
        for (uint32_t j = 0; j < C; j++) {
            c10[j] = (!w.nrow(j) && w.base2[j] < 4) ? kACTGN[w.base2[j]] : 'N';
            if (c10[j] == 'N') nidx.push_back(j), c10[j] = 'A';
        }
        std::vector<uint32_t> code = encode2(c10);
        const uint64_t param[3] = {nidx.size(), code.size(), C};
        put(f10_, param, 8, 3);
        put(f10_, code.data(), 4, code.size());
        put(f10_, nidx.data(), 4, nidx.size());
    }
    // columns 11, 12, 13 (cpp/compress.h:951-1175); col 13 is only kept
    // through the common list or when col 11 and col 12 are both zero
    put(f11_, &nind, 4, 1);
    if (!all_n) {
        std::vector<int32_t> cid, c11, c12, c13, x12i, x12v, x13i, x13v;
        for (uint32_t j = 0; j < C; j++) {
            int32_t a = 0, b = 0, d = 0;
            if (!w.nrow(j) && w.base2[j] < 4) {
                const uint32_t b2 = w.base2[j];
                const int qsb = w.qs[4ull * j + b2];
                b = w.cu[4ull * j + b2];
                a = qsb == 0 ? 0 : qsb / b;
                d = w.ca[4ull * j + b2];
            }
            if (a != 0) {
                cid.push_back((int32_t)j), c11.push_back(a), c12.push_back(b), c13.push_back(d);
            } else if (a != b) {
                x12i.push_back((int32_t)j), x12v.push_back(b);
            } else if (a != d) {
                x13i.push_back((int32_t)j), x13v.push_back(d);
            }
        }
        const uint64_t param[4] = {C, cid.size(), x12i.size(), x13i.size()};
        put(f11_, param, 8, 4);
        put(f11_, x12i.data(), 4, x12i.size());
        put(f11_, x12v.data(), 4, x12v.size());
        put(f11_, x13i.data(), 4, x13i.size());
        put(f11_, x13v.data(), 4, x13v.size());
        put(f11_, cid.data(), 4, cid.size());
        for (const std::vector<int32_t>* col : {&c11, &c12, &c13}) {
            std::vector<uint32_t> words;
            std::vector<int32_t> dict;
            const uint64_t nw = dict_encode(col->data(), col->size(), words, dict);
            const uint64_t dp[2] = {nw, dict.size()};
            put(f11_, dp, 8, 2);
            put(f11_, words.data(), 4, nw);
            put(f11_, dict.data(), 4, dict.size());
        }
    }
    // columns 15 and 17 (cpp/compress.h:1178-1312)
    put(f15_, &nind, 4, 1);
    if (!all_n) {
        std::vector<int32_t> id;
        std::vector<double> val;
        for (uint32_t j = 0; j < C; j++)
            if (!w.nrow(j) && w.rank[j] != 1.0) id.push_back((int32_t)j), val.push_back(w.rank[j]);
        int32_t n = (int32_t)id.size();
        put(f15_, &n, 4, 1);
        put(f15_, id.data(), 4, id.size());
        put(f15_, val.data(), 8, val.size());
        id.clear();
        for (uint32_t j = 0; j < C; j++)
            if (!w.nrow(j) && (w.ori[j] & 8)) id.push_back((int32_t)j);
        n = (int32_t)id.size();
        put(f15_, &n, 4, 1);
        put(f15_, id.data(), 4, id.size());
    }
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

int decompress_files(const std::string& prefix, const std::string& out, std::string& err)
{
    FILE* f12 = std::fopen((prefix + ".1_2").c_str(), "r");
    if (!f12) {
        err = "cannot open " + prefix + ".1_2";
        return 1;
    }
    unsigned long nwin = 0;
    if (std::fscanf(f12, "%lu", &nwin) != 1) {
        std::fclose(f12);
        err = "bad window index";
        return 2;
    }
    FILE* f346 = std::fopen((prefix + ".3_4_6").c_str(), "rb");
    FILE* f5 = std::fopen((prefix + ".5_7_8_9_14_16").c_str(), "rb");
    FILE* f10 = std::fopen((prefix + ".10").c_str(), "rb");
    FILE* f11 = std::fopen((prefix + ".11_12_13").c_str(), "rb");
    FILE* f15 = std::fopen((prefix + ".15_17").c_str(), "rb");
    FILE* fo = std::fopen(out.c_str(), "w");
    int rc = 0;
    if (!f346 || !f5 || !f10 || !f11 || !f15 || !fo) {
        err = "cannot open the -C files";
        rc = 1;
    }
    char name[1024];
    std::vector<char> c3, c4, c6, c10;
    std::vector<int32_t> c5, c7, c8, c9, c14, c11, c12, c13, c17;
    std::vector<double> c15, c16;
    std::vector<char> line;
    for (unsigned long wi = 0; !rc && wi < nwin; wi++) {
        unsigned long ps, C;
        if (std::fscanf(f12, "%1023s\t%lu\t%lu\n", name, &ps, &C) != 3) {
            err = "bad window index row";
            rc = 2;
            break;
        }
        c3.assign(C, 'N'), c4.assign(C, 'N'), c6.assign(C, 'N'), c10.assign(C, 'N');
        c5.assign(C, 0), c7.assign(C, 0), c8.assign(C, 0), c9.assign(C, 0), c14.assign(C, 0);
        c11.assign(C, 0), c12.assign(C, 0), c13.assign(C, 0), c17.assign(C, 0);
        c15.assign(C, 1.0), c16.assign(C, 255.0);
        uint64_t pa[7];
        bool ok = get(f346, pa, 8, 7);
        if (ok && pa[0] != 1) {
            std::vector<int32_t> id4(pa[5]), id6(pa[6]), idn(pa[4]);
            std::vector<char> v4(pa[5]), v6(pa[6]);
            std::vector<uint32_t> code(pa[3]);
            ok = get(f346, id4.data(), 4, pa[5]) && get(f346, v4.data(), 1, pa[5]) && get(f346, id6.data(), 4, pa[6]) &&
                 get(f346, v6.data(), 1, pa[6]) && get(f346, idn.data(), 4, pa[4]) && get(f346, code.data(), 4, pa[3]);
            if (ok) {
                decode2(code.data(), pa[3], C, c3.data());
                for (int32_t i : idn) c3[i] = 'N';
                c4 = c3;
                for (size_t i = 0; i < id4.size(); i++) c4[id4[i]] = v4[i];
                c6 = c3;
                for (size_t i = 0; i < id6.size(); i++) c6[id6[i]] = v6[i];
            }
        }
        int32_t nind = 1;
        ok = ok && get(f5, &nind, 4, 1);
        if (ok && nind != 1)
            ok = read_rle_column(f5, C, c5.data()) && read_rle_column(f5, C, c7.data()) &&
                 read_rle_column(f5, C, c8.data()) && read_rle_column(f5, C, c9.data()) &&
                 read_rle_column(f5, C, c14.data()) && read_rle_column(f5, C, c16.data());
        ok = ok && get(f10, &nind, 4, 1);
        if (ok && nind != 1) {
            uint64_t p3[3];
            ok = get(f10, p3, 8, 3);
            std::vector<uint32_t> code(ok ? p3[1] : 0), ni(ok ? p3[0] : 0);
            ok = ok && get(f10, code.data(), 4, p3[1]) && get(f10, ni.data(), 4, p3[0]);
            if (ok) {
                decode2(code.data(), p3[1], C, c10.data());
                for (uint32_t i : ni) c10[i] = 'N';
            }
        }
        ok = ok && get(f11, &nind, 4, 1);
        if (ok && nind != 1) {
            uint64_t p4[4];
            ok = get(f11, p4, 8, 4);
            const uint64_t nc = ok ? p4[1] : 0, n12 = ok ? p4[2] : 0, n13 = ok ? p4[3] : 0;
            std::vector<int32_t> x12i(n12), x12v(n12), x13i(n13), x13v(n13), cid(nc);
            ok = ok && get(f11, x12i.data(), 4, n12) && get(f11, x12v.data(), 4, n12) && get(f11, x13i.data(), 4, n13) &&
                 get(f11, x13v.data(), 4, n13) && get(f11, cid.data(), 4, nc);
            std::vector<int32_t> vals[3];
            for (int k = 0; k < 3 && ok; k++) {
                uint64_t dp[2];
                ok = get(f11, dp, 8, 2);
                std::vector<uint32_t> words(ok ? dp[0] : 0);
                std::vector<int32_t> dict(ok ? dp[1] : 0);
                ok = ok && get(f11, words.data(), 4, dp[0]) && get(f11, dict.data(), 4, dp[1]);
                vals[k].assign(nc, 0);
                if (ok) dict_decode(words.data(), dp[0], dict.data(), dp[1], nc, vals[k].data());
            }
            if (ok) {
                for (uint64_t i = 0; i < nc; i++) c11[cid[i]] = vals[0][i], c12[cid[i]] = vals[1][i], c13[cid[i]] = vals[2][i];
                for (uint64_t i = 0; i < n12; i++) c12[x12i[i]] = x12v[i];
                for (uint64_t i = 0; i < n13; i++) c13[x13i[i]] = x13v[i];
            }
        }
        ok = ok && get(f15, &nind, 4, 1);
        if (ok && nind != 1) {
            int32_t n = 0;
            ok = get(f15, &n, 4, 1);
            std::vector<int32_t> id(ok ? n : 0);
            std::vector<double> val(ok ? n : 0);
            ok = ok && get(f15, id.data(), 4, n) && get(f15, val.data(), 8, n);
            if (ok)
                for (int32_t i = 0; i < n; i++) c15[id[i]] = val[i];
            ok = ok && get(f15, &n, 4, 1);
            if (ok) {
                id.assign(n, 0);
                ok = get(f15, id.data(), 4, n);
                if (ok)
                    for (int32_t i = 0; i < n; i++) c17[id[i]] = 1;
            }
        }
        if (!ok) {
            err = "truncated -C column file";
            rc = 2;
            break;
        }
        for (unsigned long i = 0; i < C; i++) {
            if (c3[i] == 'N' && c4[i] == 'N' && c6[i] == 'N')
                std::fprintf(fo, "%s\t%lu\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n", name, ps + i);
            else
                std::fprintf(fo, "%s\t%lu\t%c\t%c\t%d\t%c\t%d\t%d\t%d\t%c\t%d\t%d\t%d\t%d\t%#g\t%#g\t%d\n", name, ps + i,
                             c3[i], c4[i], c5[i], c6[i], c7[i], c8[i], c9[i], c10[i], c11[i], c12[i], c13[i], c14[i],
                             c15[i], c16[i], c17[i]);
        }
    }
    std::fclose(f12);
    for (FILE* f : {f346, f5, f10, f11, f15, fo})
        if (f) std::fclose(f);
    return rc;
}

}  // namespace gsnp
