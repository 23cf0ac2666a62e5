// walk.h -- dependency-adjusted quality of the sorted records of one site
// (kernel/likelihood_kernel.cu:168-256).  Shared by the per-window
// likelihood kernel and the adjoint's record keys.
#pragma once

#include <hip/hip_runtime.h>

#include "gsnp_common.h"
#include "kernels.h"

namespace gsnp {

// Walks row[0, n) in sorted order.  Within a base segment, t is the
// occurrence count of the record's (strand, cycle) and g the number of
// distinct (strand, cycle) seen so far minus one; the adjusted quality comes
// from the host-built table.  `seen` is a per-thread 16-word (512-bit) set of
// the (strand, cycle) keys of the segment; only a repeated key rescans the
// segment to count its earlier occurrences.  emit(qa, cycle, o) is called once
// per occurrence, in record order.  Returns the error flags met.
template <class F>
__device__ __forceinline__ uint32_t walk_adjusted(const uint32_t* row, uint32_t n, const Tables& t, uint32_t* seen,
                                                  F emit)
{
    int last_o = -1, gdc = -1;
    uint32_t seg = 0, bad = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t w = row[i];
        const uint32_t o = rec_base(w);
        uint32_t q = rec_q(w);
        const uint32_t cyc = rec_cycle(w), occ = rec_occ(w), key = rec_key(w);
        if ((int)o != last_o) {
#pragma unroll
            for (int k = 0; k < 16; k++) seen[k] = 0;
            gdc = -1;
            seg = i;
            last_o = (int)o;
        }
        const uint32_t word = seen[key >> 5], bit = 1u << (key & 31);
        uint32_t prior = 0;
        if (word & bit) {
            for (uint32_t j = seg; j < i; j++)
                if (rec_key(row[j]) == key) prior += rec_occ(row[j]);
        } else {
            seen[key >> 5] = word | bit;
        }
        if (q > 63) {
            bad |= kErrQadjRange;
            q = 63;
        }
        for (uint32_t k = 0; k < occ; k++) {
            if (prior + k == 0) gdc += 1;
            const uint32_t tm1 = prior + k < (uint32_t)t.td ? prior + k : (uint32_t)t.td - 1;
            const uint32_t gi = (uint32_t)gdc < (uint32_t)t.gd ? (uint32_t)gdc : (uint32_t)t.gd - 1;
            uint32_t qa = t.qadj[((size_t)q * t.td + tm1) * t.gd + gi];
            if (qa == 0) {
                bad |= kErrQadjRange;
                qa = 1;
            }
            emit(qa, cyc, o);
        }
    }
    return bad;
}

}  // namespace gsnp
