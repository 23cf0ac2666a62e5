// cols.h -- the -C column files of a called span, encoded on the device.
#pragma once

#include <stdint.h>

#include <vector>

namespace gsnp {

struct Ctx;

// The chunks of a span's windows, in window order: the bytes each of the five
// column files receives (.3_4_6, .5_7_8_9_14_16, .10, .11_12_13, .15_17) and
// the start (1-based) and size of every window for the index file .1_2.
struct ColumnChunks {
    std::vector<uint8_t> file[5];
    std::vector<uint64_t> starts;
    std::vector<uint32_t> sizes;
};

// Encodes the called sites [offset, offset+n) of the context, which hold the
// chromosome positions [ws_abs, ws_abs+n) (ws_abs on a window boundary, n > 0),
// one chunk per reference window, and appends them to `out`.
int encode_columns(Ctx& c, uint32_t offset, uint32_t ws_abs, uint32_t n, ColumnChunks& out);

}  // namespace gsnp
