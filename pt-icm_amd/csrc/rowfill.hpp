// Row fill of the bit-set cluster growth on periodic lattices.  Plain C++
// with clang builtins only, so that the same function compiles into the
// kernels (device, inlined) and into a host unit test that checks it
// against a site-by-site closure.
#pragma once

#include <cstdint>

#ifndef PTX_ROWFILL_FN
#define PTX_ROWFILL_FN __device__ __forceinline__
#endif

namespace ptx {

// x closure of one lattice row of L sites in RR words (bits >= L zero),
// wrapped at L: v (seeds, a subset of d) becomes the union of the runs of d
// that hold a seed.  Upward within each run by one addition, a & ~(a + s) |
// s (the carry from a run's lowest seed clears the run up to its top; a
// second seed in the run is put back by | s), downward the same on the
// bit-reversed field, then once more if the run through the wrap joins a
// filled half: O(RR) instructions instead of one site per iteration.
template <int RR>
PTX_ROWFILL_FN void fill_row(uint32_t* v, const uint32_t* d, int L) {
    const int sh = 32 * RR - L;      // 0..31: the reversed field's offset
    const int lb = L - 32 * (RR - 1);  // sites in the last word
    auto up_fill = [&](const uint32_t* a, const uint32_t* s, uint32_t* f) {
        uint32_t carry = 0u;
#pragma unroll
        for (int k = 0; k < RR; ++k) {
            const unsigned long long t = (unsigned long long)a[k] + s[k] + carry;
            carry = (uint32_t)(t >> 32);
            f[k] = (a[k] & ~(uint32_t)t) | s[k];
        }
    };
    auto rev = [&](const uint32_t* x, uint32_t* y) {  // bit i of the field -> bit L - 1 - i
        uint32_t t[RR];
#pragma unroll
        for (int k = 0; k < RR; ++k) t[k] = __builtin_bitreverse32(x[RR - 1 - k]);
#pragma unroll
        for (int k = 0; k < RR; ++k) {
            const uint32_t hi = (k + 1 < RR && sh != 0) ? (t[k + 1] << (32 - sh)) : 0u;
            y[k] = (sh != 0 ? (t[k] >> sh) : t[k]) | hi;
        }
    };
    uint32_t dr[RR];
    rev(d, dr);
    for (int pass = 0; pass < 2; ++pass) {
        uint32_t fu[RR], sr[RR], fdr[RR], fd[RR];
        up_fill(d, v, fu);
        rev(v, sr);
        up_fill(dr, sr, fdr);
        rev(fdr, fd);
#pragma unroll
        for (int k = 0; k < RR; ++k) v[k] = fu[k] | fd[k];
        // the wrap: sites L - 1 and 0 are neighbours
        const bool top = ((v[RR - 1] >> (lb - 1)) & 1u) != 0u, bot = (v[0] & 1u) != 0u;
        const bool dtop = ((d[RR - 1] >> (lb - 1)) & 1u) != 0u, dbot = (d[0] & 1u) != 0u;
        bool more = false;
        if (top && dbot && !bot) {
            v[0] |= 1u;
            more = true;
        }
        if (bot && dtop && !top) {
            v[RR - 1] |= 1u << (lb - 1);
            more = true;
        }
        if (!more) break;
    }
}

}  // namespace ptx
