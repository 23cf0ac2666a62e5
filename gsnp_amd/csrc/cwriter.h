// cwriter.h -- GSNP's -C compressed column format (cpp/compress.h:33-1312):
// per-window chunks in <prefix>.3_4_6, .5_7_8_9_14_16, .10, .11_12_13,
// .15_17 and the window index <prefix>.1_2.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "cols.h"

namespace gsnp {

// Bits per dictionary index, ceil(log(n)/log(2)) with a minimum of 1
// (cpp/dict_compression.h:47-54, evaluated with the host libm).
int dict_bits(uint64_t n);

// Every call returns 0 or 1 (a file could not be created or written, the
// message in err).
class CWriter {
public:
    ~CWriter();
    int open(const std::string& prefix, std::string& err);
    // the all-N chunk of an empty window
    int write_empty_window(const std::string& name, uint32_t ws, std::string& err);
    // chunks encoded on the device (encode_columns), appended in window order
    int write_chunks(const std::string& name, const ColumnPass& ch, std::string& err);
    int close(std::string& err);

private:
    std::string prefix_;
    FILE *f346_ = nullptr, *f5_ = nullptr, *f10_ = nullptr, *f11_ = nullptr, *f15_ = nullptr;
    std::vector<std::string> names_;
    std::vector<uint64_t> starts_, sizes_;
};

// One window of -C output as its 17 columns (CNSDataWindow,
// cpp/decompression.cpp:7-128): chr, 1-based start, rows; columns 3, 4, 6,
// 10 as characters, 15 and 16 as doubles, the rest as integers.
struct CnsWindow {
    std::string chr;
    uint64_t start = 0, n = 0;
    std::vector<char> c3, c4, c6, c10;
    std::vector<int32_t> c5, c7, c8, c9, c11, c12, c13, c14, c17;
    std::vector<double> c15, c16;
};

// One window's chunks as stored, every count and index checked against the
// window size: what the host decoder and the device decoder take.
struct RawRle {
    uint64_t nrun = 0, nd_val = 0, nd_len = 0;
    std::vector<uint32_t> wv, wl, dl;  // value words, length words, length dictionary
    std::vector<int32_t> dv32;         // value dictionary (columns 5, 7, 8, 9, 14)
    std::vector<double> dv64;          // value dictionary (column 16)
};
struct RawChunk {
    std::string chr;
    uint64_t start = 0, n = 0;
    bool n346 = true, n5 = true, n10 = true, n11 = true, n15 = true;  // per file: an all-N chunk
    std::vector<int32_t> id4, id6, idn;
    std::vector<char> v4, v6;
    std::vector<uint32_t> code3;
    RawRle rle[6];  // columns 5, 7, 8, 9, 14, 16
    std::vector<uint32_t> code10, ni10;
    std::vector<int32_t> x12i, x12v, x13i, x13v, cid;
    uint64_t nd11[3] = {0, 0, 0};
    std::vector<uint32_t> w11[3];
    std::vector<int32_t> d11[3];
    std::vector<int32_t> id15, id17;
    std::vector<double> v15;
};

// Reads -C files one window at a time (decompressAll's loop,
// cpp/decompression.cpp:130-255).
class ColumnReader {
public:
    ~ColumnReader();
    int open(const std::string& prefix, std::string& err);
    // got = false at the end of the files
    int next(CnsWindow& w, bool& got, std::string& err);
    // the next window's chunks, validated but not decoded
    int read_chunk(RawChunk& c, bool& got, std::string& err);

private:
    FILE *f12_ = nullptr, *f346_ = nullptr, *f5_ = nullptr, *f10_ = nullptr, *f11_ = nullptr, *f15_ = nullptr;
    unsigned long nwin_ = 0, wi_ = 0;
};

// A validated chunk decoded into the 17 columns on the host.
int decode_chunk(const RawChunk& c, CnsWindow& w, std::string& err);

// -C files back to .cns text (cpp/decompression.cpp:130-255).
int decompress_files(const std::string& prefix, const std::string& out, std::string& err);

// Bits per dictionary index and words of n indices (cpp/dict_compression.h:47-58).
uint64_t dict_words(uint64_t n, uint64_t nd);

}  // namespace gsnp
