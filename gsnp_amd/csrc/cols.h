// cols.h -- the -C column files of a called span, encoded on the device.
#pragma once

#include <stdint.h>

#include <functional>
#include <vector>

namespace gsnp {

struct Ctx;

// One encoder pass (up to 32 windows) of a span: the bytes each of the five
// column files receives (.3_4_6, .5_7_8_9_14_16, .10, .11_12_13, .15_17), in
// window order, in pinned host memory that the next pass reuses, and the
// start (1-based) and size of every window for the index file .1_2.
struct ColumnPass {
    const uint8_t* file[5];
    uint64_t bytes[5];
    const uint64_t* starts;
    const uint32_t* sizes;
    const uint64_t* offsets;  // [n_windows][5]: where window w's chunk starts in file f's bytes
    uint32_t n_windows;
};
using ColumnSink = std::function<int(const ColumnPass&)>;

// A queued pass whose copy to pinned host memory may still be in flight, and
// the sink it goes to.
struct ColumnPending {
    ColumnPass pass{};
    int buf = 0;
    std::vector<uint64_t> starts;
    std::vector<uint32_t> sizes;
    std::vector<uint64_t> offsets;
    ColumnSink sink;
};

// Encodes the called sites [offset, offset+n) of the context, which hold the
// chromosome positions [ws_abs, ws_abs+n) (ws_abs on a window boundary, n > 0),
// one chunk per reference window, and hands every pass to `sink` in order.
// The last pass stays queued, its copy running beside whatever the caller
// queues next, until the next call hands it over or finish_columns does: a
// caller calls finish_columns before it writes anything else to the same
// files and before its sink's captures go out of scope, or drop_columns after
// an error.
int encode_columns(Ctx& c, uint32_t offset, uint32_t ws_abs, uint32_t n, const ColumnSink& sink);
int finish_columns(Ctx& c);
void drop_columns(Ctx& c);

}  // namespace gsnp
