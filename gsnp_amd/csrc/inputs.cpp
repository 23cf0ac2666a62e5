// inputs.cpp -- FASTA, known-SNP and SOAP alignment readers (host).
#include "inputs.h"

#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <system_error>
#include <thread>

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace gsnp {

namespace {

// Buffered line reader over a FILE*.
class LineReader {
public:
    explicit LineReader(FILE* f) : f_(f), buf_(1 << 22) {}
    // returns false at end of file; line excludes the '\n'
    bool next(const char*& s, size_t& n)
    {
        for (;;) {
            char* nl = (char*)std::memchr(buf_.data() + beg_, '\n', end_ - beg_);
            if (nl) {
                s = buf_.data() + beg_;
                n = (size_t)(nl - s);
                beg_ = (size_t)(nl - buf_.data()) + 1;
                return true;
            }
            if (eof_) {
                if (beg_ < end_) {
                    s = buf_.data() + beg_;
                    n = end_ - beg_;
                    beg_ = end_;
                    return true;
                }
                return false;
            }
            if (beg_ > 0) {
                std::memmove(buf_.data(), buf_.data() + beg_, end_ - beg_);
                end_ -= beg_;
                beg_ = 0;
            }
            if (end_ == buf_.size()) buf_.resize(buf_.size() * 2);
            size_t got = std::fread(buf_.data() + end_, 1, buf_.size() - end_, f_);
            if (got == 0) eof_ = true;
            end_ += got;
        }
    }

private:
    FILE* f_;
    std::vector<char> buf_;
    size_t beg_ = 0, end_ = 0;
    bool eof_ = false;
};

struct FileCloser {
    void operator()(FILE* f) const
    {
        if (f) std::fclose(f);
    }
};
using FilePtr = std::unique_ptr<FILE, FileCloser>;

inline bool acgt(char c)
{
    switch (c) {
    case 'A': case 'C': case 'G': case 'T': case 'a': case 'c': case 'g': case 't':
        return true;
    default:
        return false;
    }
}

// Token scanner with std::istream extraction semantics for the SOAP fields.
struct Scan {
    const char* p;
    const char* e;
    void ws()
    {
        while (p < e && std::isspace((unsigned char)*p)) p++;
    }
    bool word(const char*& s, int& n)
    {
        ws();
        if (p >= e) return false;
        s = p;
        while (p < e && !std::isspace((unsigned char)*p)) p++;
        n = (int)(p - s);
        return true;
    }
    bool integer(int& v)
    {
        ws();
        const char* q = p;
        bool neg = false;
        if (q < e && (*q == '+' || *q == '-')) neg = *q++ == '-';
        if (q >= e || !std::isdigit((unsigned char)*q)) return false;
        long long x = 0;
        while (q < e && std::isdigit((unsigned char)*q)) {
            x = x * 10 + (*q++ - '0');
            if (x > 2147483648LL) return false;
        }
        if (neg) x = -x;
        if (x > INT_MAX || x < INT_MIN) return false;
        v = (int)x;
        p = q;
        return true;
    }
    bool character(char& c)
    {
        ws();
        if (p >= e) return false;
        c = *p++;
        return true;
    }
};

}  // namespace

int load_genome(const char* fasta, const char* dbsnp, Genome& g, std::string& err)
{
    FilePtr f(std::fopen(fasta, "r"));
    if (!f) {
        err = std::string("cannot open ") + fasta;
        return GSNP_E_IO;
    }
    LineReader lr(f.get());
    const char* s;
    size_t n;
    int cur = -1;
    while (lr.next(s, n)) {
        if (n && s[0] == '>') {
            size_t i = 1;
            while (i < n && !std::isspace((unsigned char)s[i])) i++;
            std::string name(s + 1, i - 1);
            if (g.index.count(name)) {
                err = "duplicate chromosome " + name;
                return GSNP_E_FORMAT;
            }
            g.index[name] = (int)g.chr.size();
            g.chr.emplace_back();
            g.chr.back().name = name;
            cur = (int)g.chr.size() - 1;
        } else if (cur >= 0) {
            auto& code = g.chr[cur].code;
            size_t at = code.size();
            code.resize(at + n);
            for (size_t k = 0; k < n; k++) code[at + k] = (uint8_t)(((unsigned)(signed char)s[k] >> 1) & 7);
        }
    }
    if (dbsnp && dbsnp[0]) {
        FilePtr fs(std::fopen(dbsnp, "r"));
        if (!fs) {
            err = std::string("cannot open ") + dbsnp;
            return GSNP_E_IO;
        }
        LineReader ls(fs.get());
        struct Tmp {
            uint32_t pos, order;
            gsnp_snp s;
        };
        std::vector<std::vector<Tmp>> per(g.chr.size());
        uint32_t order = 0;
        while (ls.next(s, n)) {
            std::string line(s, n);
            char name[256];
            long long pos;
            int hm, va, in;
            double fr[4];
            if (std::sscanf(line.c_str(), "%255s %lld %d %d %d %lf %lf %lf %lf", name, &pos, &hm, &va, &in, &fr[0],
                            &fr[1], &fr[2], &fr[3]) != 9)
                continue;
            auto it = g.index.find(name);
            if (it == g.index.end()) continue;
            Chrom& c = g.chr[it->second];
            if (pos < 1 || (uint64_t)pos > c.code.size()) {
                err = "known-SNP position outside chromosome " + c.name;
                return GSNP_E_RANGE;
            }
            Tmp t;
            t.pos = (uint32_t)(pos - 1);
            t.order = order++;
            std::memset(&t.s, 0, sizeof t.s);
            t.s.hapmap = hm != 0;
            t.s.validated = va != 0;
            t.s.indel = in != 0;
            std::memcpy(t.s.freq, fr, sizeof fr);
            per[it->second].push_back(t);
        }
        for (size_t ci = 0; ci < g.chr.size(); ci++) {
            auto& v = per[ci];
            std::stable_sort(v.begin(), v.end(), [](const Tmp& a, const Tmp& b) { return a.pos < b.pos; });
            Chrom& c = g.chr[ci];
            for (const Tmp& t : v) {
                if (!c.snp_pos.empty() && c.snp_pos.back() == t.pos) continue;  // first insertion wins
                c.snp_pos.push_back(t.pos);
                c.snp.push_back(t.s);
                c.code[t.pos] |= 8;
            }
        }
    }
    return GSNP_OK;
}

namespace {

// One SOAP line that passed the per-line checks, in file order.
struct ParsedRead {
    int32_t ci, pos, len_hit;
    char strand;
    bool beyond;  // ends at or past the chromosome end: kept out of the reads
};

// The lines of one slice of the file: its reads (seq/qual `L` bytes each)
// up to its first per-line error, and that error.
struct ParsedSlice {
    std::vector<ParsedRead> reads;
    std::vector<char> seq, qual;
    int err = GSNP_OK;
    std::string msg;
};

// Per-line part of the SOAP reader (cpp/soap_snp.h:132-169): fields, indel
// edits, chromosome lookup and range checks.  What depends on earlier lines
// (chromosome grouping, the order check) is left to the merge.
void parse_slice(const char* b, const char* e, const gsnp_params& p, const Genome& g, ParsedSlice& out)
{
    const int L = p.read_length;
    std::vector<char> rd(1024), ql(1024);
    std::string name;
    while (b < e) {
        const char* nl = static_cast<const char*>(std::memchr(b, '\n', (size_t)(e - b)));
        const char* le = nl ? nl : e;
        Scan sc{b, le};
        b = nl ? nl + 1 : e;
        const char *id, *r0, *q0, *chr;
        int nid, nr, nq, nchr, hit, rlen, pos, mm;
        char ab, strand;
        if (!sc.word(id, nid) || !sc.word(r0, nr) || !sc.word(q0, nq) || !sc.integer(hit) || !sc.character(ab) ||
            !sc.integer(rlen) || !sc.character(strand) || !sc.word(chr, nchr) || !sc.integer(pos) || !sc.integer(mm))
            continue;
        // indel edits (cpp/soap_snp.h:139-158); the read keeps its length field
        size_t cap = (size_t)nr + (size_t)nq + 2048;
        if (rd.size() < cap) rd.resize(cap), ql.resize(cap);
        int elen = nr, eqlen = nq;
        const char* es = r0;
        const char* eq = q0;
        if (mm > 100) {
            int ip;
            if (!sc.integer(ip)) continue;
            int il = mm > 200 ? mm - 200 : mm - 100;
            bool ins_n = mm > 200;
            bool bad = ins_n ? (ip < 0 || ip > nr || ip > nq || il > 1024) : (ip < 0 || ip + il > nr || ip + il > nq);
            if (bad) {
                out.err = GSNP_E_FORMAT;
                out.msg = "indel position out of range";
                return;
            }
            int skip = ins_n ? 0 : il;
            int tail = rlen - ip - skip;
            if (tail < 0) tail = 0;
            int tr = std::min(nr - ip - skip, tail), tq = std::min(nq - ip - skip, tail);
            std::memcpy(rd.data(), r0, ip);
            std::memcpy(ql.data(), q0, ip);
            int at = ip;
            if (ins_n) {
                std::memset(rd.data() + at, 'N', il);
                std::memset(ql.data() + at, 'N', il);
                at += il;
            }
            std::memcpy(rd.data() + at, r0 + ip + skip, tr);
            std::memcpy(ql.data() + at, q0 + ip + skip, tq);
            elen = at + tr;
            eqlen = at + tq;
            es = rd.data();
            eq = ql.data();
        }
        pos -= 1;
        if (pos < 0) continue;
        name.assign(chr, nchr);
        auto it = g.index.find(name);
        if (it == g.index.end()) {
            out.err = GSNP_E_CHROM;
            out.msg = "Chromosome: !" + name + "! NOT found";
            return;
        }
        if (rlen < 1 || rlen > L) {
            out.err = GSNP_E_RANGE;
            out.msg = "read length " + std::to_string(rlen) + " outside [1," + std::to_string(L) + "]";
            return;
        }
        if (hit < 0 || hit > 65535) {
            out.err = GSNP_E_RANGE;
            out.msg = "hit count " + std::to_string(hit) + " outside [0,65535]";
            return;
        }
        if (hit == 1) {
            for (int k = 0; k < rlen && k < elen; k++) {
                int qc = k < eqlen ? (unsigned char)eq[k] : 0;
                if (!acgt(es[k]) || qc < p.q_min) continue;
                int q = qc - p.q_min;
                if (q < 1 || q > 63) {
                    out.err = GSNP_E_QUAL;
                    out.msg = "quality " + std::to_string(q) + " outside [1,63] on a unique-hit base";
                    return;
                }
            }
        }
        const int ci = it->second;
        ParsedRead r;
        r.ci = ci;
        r.pos = pos;
        r.len_hit = (int32_t)(((uint32_t)rlen << 16) | (uint32_t)hit);
        r.strand = strand;
        r.beyond = (int64_t)pos + rlen >= (int64_t)g.chr[ci].code.size();
        out.reads.push_back(r);
        if (r.beyond) continue;
        const size_t at = out.seq.size();
        out.seq.resize(at + L);
        out.qual.resize(at + L);
        for (int k = 0; k < L; k++) {
            out.seq[at + k] = k < elen ? es[k] : 'N';
            out.qual[at + k] = k < eqlen ? eq[k] : 0;
        }
    }
}

// The file's bytes, memory-mapped (or read, where mapping is refused).
struct FileBytes {
    const char* p = nullptr;
    size_t n = 0;
    void* map = nullptr;
    std::vector<char> copy;
    ~FileBytes()
    {
        if (map) munmap(map, n);
    }
    bool open(const char* path)
    {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) == 0 && S_ISREG(st.st_mode) && st.st_size > 0) {
            void* m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m != MAP_FAILED) {
                map = m;
                n = (size_t)st.st_size;
                p = static_cast<const char*>(m);
                ::close(fd);
                return true;
            }
        }
        char buf[1 << 16];
        ssize_t k;
        while ((k = ::read(fd, buf, sizeof buf)) > 0) copy.insert(copy.end(), buf, buf + k);
        ::close(fd);
        if (k < 0) return false;
        p = copy.data();
        n = copy.size();
        return true;
    }
};

}  // namespace

int load_alignments(const char* soap, const gsnp_params& p, const Genome& g, std::vector<ReadBlock>& blocks,
                    std::string& err)
{
    FileBytes fb;
    if (!fb.open(soap)) {
        err = std::string("cannot open ") + soap;
        return GSNP_E_IO;
    }
    const int L = p.read_length;
    // slices of whole lines, parsed on their own threads (up to 16, one per
    // 4 MB; GSNP_PARSE_THREADS sets the number of slices)
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    size_t T = std::max<size_t>(1, std::min<size_t>({hw, 16, fb.n / (4u << 20) + 1}));
    if (const char* v = std::getenv("GSNP_PARSE_THREADS")) T = std::max<size_t>(1, std::min<size_t>(64, std::strtoul(v, nullptr, 10)));
    std::vector<size_t> cut(T + 1, fb.n);
    cut[0] = 0;
    for (size_t i = 1; i < T; i++) {
        size_t c = std::max(cut[i - 1], fb.n * i / T);
        const char* nl = c < fb.n ? static_cast<const char*>(std::memchr(fb.p + c, '\n', fb.n - c)) : nullptr;
        cut[i] = nl ? (size_t)(nl - fb.p) + 1 : fb.n;
    }
    std::vector<ParsedSlice> sl(T);
    {
        // a slice whose thread cannot be started is parsed here afterwards
        std::vector<std::thread> th;
        std::vector<char> started(T, 0);
        for (size_t i = 1; i < T; i++) {
            try {
                th.emplace_back([&, i] { parse_slice(fb.p + cut[i], fb.p + cut[i + 1], p, g, sl[i]); });
                started[i] = 1;
            } catch (const std::system_error&) {
            }
        }
        parse_slice(fb.p + cut[0], fb.p + cut[1], p, g, sl[0]);
        for (auto& t : th) t.join();
        for (size_t i = 1; i < T; i++)
            if (!started[i]) parse_slice(fb.p + cut[i], fb.p + cut[i + 1], p, g, sl[i]);
    }
    // kept reads per block, so that each block is allocated once
    std::vector<uint64_t> per_block;
    {
        int c = -1;
        for (const ParsedSlice& s : sl)
            for (const ParsedRead& r : s.reads) {
                if (r.ci != c) per_block.push_back(0), c = r.ci;
                per_block.back() += r.beyond ? 0 : 1;
            }
    }
    // merge in file order: chromosome grouping, the order check
    // (cpp/call_genotype.cc:495-519), the first error in file order
    std::vector<char> seen(g.chr.size(), 0);
    int cur = -1;
    int32_t last = 0;
    for (size_t i = 0; i < T; i++) {
        ParsedSlice& s = sl[i];
        // the slice's kept reads so far (k) and the first of them not yet
        // copied into the current block (run): runs are copied whole
        size_t k = 0, run = 0;
        auto flush = [&] {
            if (k > run) {
                ReadBlock& b = blocks.back();
                b.seq.insert(b.seq.end(), s.seq.begin() + run * L, s.seq.begin() + k * L);
                b.qual.insert(b.qual.end(), s.qual.begin() + run * L, s.qual.begin() + k * L);
            }
            run = k;
        };
        for (const ParsedRead& r : s.reads) {
            if (r.ci != cur) {
                flush();
                if (seen[r.ci]) {
                    err = "reads of chromosome " + g.chr[r.ci].name + " are not contiguous";
                    return GSNP_E_UNSORTED;
                }
                seen[r.ci] = 1;
                blocks.emplace_back();
                ReadBlock& nb = blocks.back();
                nb.chr = r.ci;
                nb.pitch = (uint32_t)L;
                const uint64_t want = per_block[blocks.size() - 1];
                nb.pos.reserve(want);
                nb.len_hit.reserve(want);
                nb.strand.reserve(want);
                nb.seq.reserve(want * L);
                nb.qual.reserve(want * L);
                cur = r.ci;
                last = 0;
            }
            if (r.beyond) continue;
            if (r.pos < last) {
                flush();
                err = "Errors in sorting:" + std::to_string(r.pos) + "<" + std::to_string(last);
                return GSNP_E_UNSORTED;
            }
            last = r.pos;
            ReadBlock& b = blocks.back();
            b.pos.push_back(r.pos);
            b.len_hit.push_back(r.len_hit);
            b.strand.push_back(r.strand);
            k++;
        }
        flush();
        if (s.err) {
            err = s.msg;
            return s.err;
        }
        std::vector<ParsedRead>().swap(s.reads);  // release the slice as the merge passes it
        std::vector<char>().swap(s.seq);
        std::vector<char>().swap(s.qual);
    }
    return GSNP_OK;
}

// The matrix from calibration counts (cpp/matrix.cc:129-205), cnt indexed
// (qc - q_min) << 12 | cycle << 4 | ref << 2 | obs for quality characters qc.
void matrix_from_counts(const uint64_t* cnt, const gsnp_params& p, double* pm)
{
    for (size_t i = 0; i < (size_t)256 * 256 * 16; i++) pm[i] = 1.0;
    const uint64_t kPow = 10;  // "statistically powerful" count
    for (int qc = p.q_min; qc <= p.q_max; qc++) {
        const uint64_t* row = &cnt[(size_t)(qc - p.q_min) << 12];
        uint64_t by_type[16] = {0}, by_ref[4] = {0}, total = 0, mism = 0;
        for (int c = 0; c < p.read_length; c++)
            for (int t = 0; t < 16; t++) {
                uint64_t v = row[(c << 4) | t];
                by_type[t] += v;
                by_ref[t >> 2] += v;
                total += v;
                if (t % 5) mism += v;
            }
        const int q = qc - p.q_min;
        for (int c = 0; c < p.read_length; c++) {
            uint64_t sum[4] = {0};
            for (int t = 0; t < 16; t++) sum[t >> 2] += row[(c << 4) | t];
            for (int r = 0; r < 4; r++)
                for (int o = 0; o < 4; o++) {
                    double& v = pm[((size_t)q << 12) | ((size_t)c << 4) | (r << 2) | o];
                    const uint64_t k = row[(c << 4) | (r << 2) | o];
                    if (k > kPow)
                        v = (double)k / (double)sum[r];
                    else if (by_type[(r << 2) | o] > kPow)
                        v = (double)by_type[(r << 2) | o] / (double)by_ref[r];
                    else if (total > 0)
                        v = (o == r) ? (double)(total - mism) / (double)total : (double)mism / (double)total;
                    if (v == 0 || v == 1) {  // fall back to the Phred value
                        double e = std::pow(10, -(q / 10.0));
                        if (o == r) {
                            v = 1 - e;
                            if (v < 0.25) v = 0.25;
                        } else {
                            v = e / 3;
                            if (v > 0.25) v = 0.25;
                        }
                    }
                }
        }
    }
}

int matrix_read(const char* path, double* pm, std::string& err)
{
    FilePtr f(std::fopen(path, "r"));
    if (!f) {
        err = std::string("cannot open ") + path;
        return GSNP_E_IO;
    }
    for (size_t i = 0; i < (size_t)256 * 256 * 16; i++) pm[i] = 1.0;
    LineReader lr(f.get());
    const char* s;
    size_t n;
    while (lr.next(s, n)) {
        std::string line(s, n);
        char* p = (char*)line.c_str();
        char* e;
        long q = std::strtol(p, &e, 10);
        if (e == p) continue;
        p = e;
        long c = std::strtol(p, &e, 10);
        if (e == p || q < 0 || q > 255 || c < 0 || c > 255) continue;
        p = e;
        for (int t = 0; t < 16; t++) {
            double v = std::strtod(p, &e);
            if (e == p) break;
            p = e;
            pm[((size_t)q << 12) | ((size_t)c << 4) | t] = v;
        }
    }
    return GSNP_OK;
}

int matrix_write(const char* path, const double* pm, const gsnp_params& p, std::string& err)
{
    FILE* f = std::fopen(path, "w");
    if (!f) {
        err = std::string("cannot create ") + path;
        return GSNP_E_IO;
    }
    bool ok = true;
    for (int q = 0; q <= p.q_max - p.q_min && ok; q++)
        for (int c = 0; c < p.read_length && ok; c++) {
            ok = std::fprintf(f, "%d\t%d", q, c) > 0;
            for (int t = 0; t < 16 && ok; t++)
                ok = std::fprintf(f, "\t%.16e", pm[((size_t)q << 12) | ((size_t)c << 4) | t]) > 0;
            ok = ok && std::fputc('\n', f) != EOF;
        }
    ok = std::fclose(f) == 0 && ok;
    if (!ok) {
        err = std::string("cannot write ") + path;
        return GSNP_E_IO;
    }
    return GSNP_OK;
}

// Genome::read_region and Chr_info::set_region (cpp/chromosome.cc:66-170):
// each line "chr start end" (1-based, inclusive) marks the windows that
// [start - read_length, end - 1], clipped to [0, length], touches.  Lines of
// unknown chromosomes are skipped, as in the reference.
int load_regions(const char* path, const Genome& g, int read_length, std::vector<std::vector<uint64_t>>& win_mask,
                 std::string& err)
{
    FilePtr f(std::fopen(path, "r"));
    if (!f) {
        err = std::string("cannot open ") + path;
        return GSNP_E_IO;
    }
    win_mask.assign(g.chr.size(), {});
    LineReader lr(f.get());
    const char* s;
    size_t n;
    while (lr.next(s, n)) {
        std::string line(s, n);
        char name[1024];
        long long a, b;
        if (std::sscanf(line.c_str(), "%1023s %lld %lld", name, &a, &b) != 3) {
            err = "Wrong format in target region file";
            return GSNP_E_FORMAT;
        }
        auto it = g.index.find(name);
        if (it == g.index.end()) continue;  // "Unexpected Chromosome"
        const uint64_t len = g.chr[it->second].code.size();
        long long lo = a - read_length, hi = b - 1;
        lo = lo < 0 ? 0 : (lo >= (long long)len ? (long long)len : lo);
        hi = hi < 0 ? 0 : (hi >= (long long)len ? (long long)len : hi);
        if (lo > hi) {
            err = "Invalid region: " + std::to_string(lo) + "-" + std::to_string(hi);
            return GSNP_E_FORMAT;
        }
        std::vector<uint64_t>& m = win_mask[it->second];
        if (m.empty()) m.assign((len / GSNP_WIN + 1) / 64 + 1, 0);
        for (uint64_t w = (uint64_t)lo / GSNP_WIN; w <= (uint64_t)hi / GSNP_WIN; w++) m[w / 64] |= 1ull << (63 - w % 64);
    }
    return GSNP_OK;
}

}  // namespace gsnp
