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

// Host copy of one called window, the inputs of every column.
struct WindowCols {
    uint32_t n = 0;
    std::vector<uint8_t> ori, base1, base2, type1;
    std::vector<int32_t> q_cns, depth, repeat;
    std::vector<int32_t> qs, cu, ca;  // 4 per site
    std::vector<double> rank;
    void resize(uint32_t c)
    {
        n = c;
        ori.resize(c), base1.resize(c), base2.resize(c), type1.resize(c);
        q_cns.resize(c), depth.resize(c), repeat.resize(c), rank.resize(c);
        qs.resize(4ull * c), cu.resize(4ull * c), ca.resize(4ull * c);
    }
    bool nrow(uint32_t j) const { return ((ori[j] & 4) && depth[j] == 0) || base1[j] >= 4; }
    bool nflag(uint32_t j) const { return (ori[j] & 4) && depth[j] == 0; }
};

// Bits per dictionary index, ceil(log(n)/log(2)) with a minimum of 1
// (cpp/dict_compression.h:47-54, evaluated with the host libm).
int dict_bits(uint64_t n);

// Sorted-unique dictionary + MSB-first packing of indices, last word
// right-aligned (cpp/dict_compression.h:116-166).  Returns the word count.
template <class T>
uint64_t dict_encode(const T* in, uint64_t n, std::vector<uint32_t>& words, std::vector<T>& dict);

class CWriter {
public:
    int open(const std::string& prefix, std::string& err);
    void write_window(const std::string& name, uint32_t ws, const WindowCols& w);
    // chunks encoded on the device (encode_columns), appended in window order
    void write_chunks(const std::string& name, const ColumnPass& ch);
    int close(std::string& err);

private:
    std::string prefix_;
    FILE *f346_ = nullptr, *f5_ = nullptr, *f10_ = nullptr, *f11_ = nullptr, *f15_ = nullptr;
    std::vector<std::string> names_;
    std::vector<uint64_t> starts_, sizes_;
};

// -C files back to .cns text (cpp/decompression.cpp:130-255).
int decompress_files(const std::string& prefix, const std::string& out, std::string& err);

}  // namespace gsnp
