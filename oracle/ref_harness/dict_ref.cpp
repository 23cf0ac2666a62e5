// Thin extern "C" wrappers that instantiate the reference's own CPU
// dictionary codec from cpp/dict_compression.h (compiled in place from the
// reference tree by oracle/Makefile, target `ref`).  Used only by tests to pin
// the oracle's and the product's -C dictionary encoding.
#include <vector>
#include <utility>
#include <cstring>
#include "dict_compression.h"

extern "C" {
int ref_num_bits(unsigned long n) { return getNumBitPerElement(n); }

unsigned long ref_dict_encode_i32(const int* in, unsigned long n, unsigned int* words, int* dict,
                                  unsigned long* n_dict)
{
    uint64 nw = n, nd = n;
    dictCompression(words, &nw, dict, &nd, in, n);
    *n_dict = nd;
    return nw;
}

unsigned long ref_dict_encode_f64(const double* in, unsigned long n, unsigned int* words, double* dict,
                                  unsigned long* n_dict)
{
    uint64 nw = n, nd = n;
    dictCompression(words, &nw, dict, &nd, in, n);
    *n_dict = nd;
    return nw;
}

unsigned long ref_dict_encode_char(const char* in, unsigned long n, unsigned int* words, char* dict,
                                   unsigned long* n_dict)
{
    uint64 nw = n, nd = n;
    dictCompression(words, &nw, dict, &nd, in, n);
    *n_dict = nd;
    return nw;
}

unsigned long ref_dict_encode_u32(const unsigned int* in, unsigned long n, unsigned int* words, unsigned int* dict,
                                  unsigned long* n_dict)
{
    uint64 nw = n, nd = n;
    dictCompression(words, &nw, dict, &nd, in, n);
    *n_dict = nd;
    return nw;
}

void ref_dict_decode_i32(int* out, unsigned long n, const unsigned int* words, unsigned long nw, const int* dict,
                         unsigned long nd)
{
    dictDecompression(out, n, words, nw, dict, nd);
}
}
