// inputs.h -- FASTA, known-SNP and SOAP alignment readers (host).
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gsnp_abi.h"

namespace gsnp {

struct Chrom {
    std::string name;
    std::vector<uint8_t> code;     // 4-bit code per base
    std::vector<uint32_t> snp_pos; // sorted, first entry per position kept
    std::vector<gsnp_snp> snp;
};

struct Genome {
    std::vector<Chrom> chr;
    std::unordered_map<std::string, int> index;
};

// cpp/chromosome.cc:172-220 (FASTA binarization (ch>>1)&7 and known SNPs).
int load_genome(const char* fasta, const char* dbsnp, Genome& g, std::string& err);

// The alignments of one chromosome that the calling path uses: pos >= 0 and
// pos + len < chromosome length, position sorted, input order kept.
struct ReadBlock {
    int chr = -1;
    uint32_t pitch = 0;
    std::vector<int32_t> pos;
    std::vector<int32_t> len_hit;  // (len << 16) | hit
    std::vector<char> strand;
    std::vector<char> seq;   // pitch chars per read
    std::vector<char> qual;
    uint64_t size() const { return pos.size(); }
};

// Parses SOAP lines (cpp/soap_snp.h:132-169), groups them by chromosome in
// order of first appearance, applies the filters and the order check
// (cpp/call_genotype.cc:495-519) and the input range checks.
int load_alignments(const char* soap, const gsnp_params& p, const Genome& g, std::vector<ReadBlock>& blocks,
                    std::string& err);

// Counting-trained calibration matrix (cpp/matrix.cc:42-213) over the blocks.
void matrix_from_counts(const uint64_t* cnt, const gsnp_params& p, double* pm);
int matrix_read(const char* path, double* pm, std::string& err);
int matrix_write(const char* path, const double* pm, const gsnp_params& p, std::string& err);

// -T target regions (cpp/chromosome.cc:66-170): per chromosome of g, a bit
// per window (window w at bit 63 - w % 64 of word w / 64) set where a region
// touches the window; empty for a chromosome without a region line.
int load_regions(const char* path, const Genome& g, int read_length, std::vector<std::vector<uint64_t>>& win_mask,
                 std::string& err);
inline bool window_called(const std::vector<uint64_t>& m, uint64_t w)
{
    return w / 64 < m.size() && ((m[w / 64] >> (63 - w % 64)) & 1);
}

}  // namespace gsnp
