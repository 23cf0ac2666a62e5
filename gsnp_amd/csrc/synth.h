// synth.h -- the seeded synthetic data sets of the benchmark configurations,
// generated in memory (and written as FASTA / SOAP / known-SNP files).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gsnp {

struct SynthRead {
    int32_t pos;  // 0-based
    uint8_t strand, hit;  // strand 1 = reverse
    char ab;
    int16_t indel;      // SOAP mismatch code > 100 (deletion) or > 200 (insertion), 0 if none
    int16_t indel_pos;
    uint32_t id;
};

struct SynthSnp {
    uint32_t pos;
    uint8_t hapmap, validated, indel;
    double freq[4];  // A C T G, at the six decimals of the written table
};

struct SynthContig {
    std::string name;
    std::vector<uint8_t> ref;  // A0 C1 T2 G3, 7 for N
    std::vector<SynthRead> reads;  // position order
    std::vector<char> seq, qual;   // read_len characters per read
    std::vector<SynthSnp> snps;    // sorted, unique positions
};

struct SynthSet {
    int read_len = 0;
    std::vector<SynthContig> contigs;
};

// Generates configuration `config` (1-6) at `scale` on `threads` threads;
// `only` (may be null) restricts it to those contig indices.
int synth_generate(int config, uint64_t seed, double scale, const std::vector<int>* only, int threads, SynthSet& out);
// contig count (and lengths) of a configuration
int synth_contig_count(int config, uint64_t seed, double scale, std::vector<uint64_t>* lens);
// FASTA, SOAP and known-SNP files; returns the alignment lines or -1
long long synth_write_files(const SynthSet& s, const std::string& prefix, int threads);
// generator threads: GSNP_SYNTH_THREADS, else the hardware threads capped by
// OMP_NUM_THREADS (at most 64)
int synth_threads();

}  // namespace gsnp
