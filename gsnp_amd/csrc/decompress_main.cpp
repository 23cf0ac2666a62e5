// gsnp-decompress -- -C column files back to .cns text (cpp/decompression.cpp:406-410).
#include <cstdio>

#include "../../include/gsnp_abi.h"

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::printf("Two parameters:\n1. the input compressed file.\n2. the output file name.\n");
        return 1;
    }
    int rc = gsnp_decompress(argv[1], argv[2]);
    if (rc) {
        std::fprintf(stderr, "gsnp-decompress: error %d: %s\n", rc, gsnp_last_error());
        return 1;
    }
    return 0;
}
