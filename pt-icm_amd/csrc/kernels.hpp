// kernels.hpp -- device-side building blocks shared by the split kernels
// (one launch per phase) and the fused per-instance kernel.
//
// A wavefront (64 lanes) owns one replica during a sweep; the replica's spins
// are expanded to one byte per spin in LDS so a neighbour read is a single
// ds_read_u8 and a flip a plain byte store (inside a colour class no two
// updated sites are adjacent, so no atomics are needed).
#pragma once

#include <hip/hip_runtime.h>

#include "ptx_common.h"
#include "rowfill.hpp"

namespace ptx {

constexpr int kWave = 64;

// Adjacency record formats.  Wide: u32 = position | (int16 weight << 16).
// Unit: u16 = position | (weight < 0) << K, for |w| == 1 and n <= 32736, with
// K = 15 or the instance's sign-table bit (InstDesc::kbit, see sign_table_bit).
// Rows start on 16-byte boundaries and take at least one 16-byte chunk;
// padding records name the sentinel position 32 * nw (weight 0 / sign 0),
// whose spin byte is kept 0, so they add nothing to a local field.
enum : int { kFmtWide = 0, kFmtUnit = 1 };
constexpr int kMaxDegree = 511;

// Arguments common to every kernel (device pointers of the engine).
struct DevArgs {
    const InstDesc* desc;
    const uint32_t* rowinfo; // per position: row start (16-byte units) << 9 | degree
    const uint8_t* adj;      // padded adjacency rows of every instance
    const int32_t* cls;      // class_start
    const int32_t* hfield;   // h by position
    const int32_t* p2o;      // pos2orig
    const uint32_t* thr;     // thresholds [nt][fmax]
    const double* beta;      // slot betas
    uint32_t* spins;         // [nrep][nw]
    uint32_t* best_spins;    // [nrep][nw]
    int32_t* E;
    int32_t* bestE;
    int32_t* best_sweep;
    unsigned long long* flips;
    int32_t* slot2rep;
    int32_t* rep2slot;
    uint32_t* swap_att;
    uint32_t* swap_succ;
    unsigned long long* swaps;  // per instance
    uint32_t* hist;
    uint32_t* inst_step;     // per-instance PT step counter (every mode; split mode advances it with k_step_advance)
    const int2* items;       // (instance, first replica) per sweep workgroup
    const int2* pairs;       // (instance, slot t) per ICM wavefront
    const int32_t* order;    // fused mode: instance per workgroup
    const uint32_t* lat;     // lattice tables (LatView), one per lattice shape
    const uint32_t* jneg;    // lattice edge signs per instance (InstDesc::jn_off)
    const uint32_t* lsign;   // edge signs of the multispin lattice sweep per instance (InstDesc::lsw_off)
    int32_t fmt;             // adjacency record format
    int32_t tab;             // 1: unit records carry the sign at bit kbit and k_sweep keeps the sign table
    unsigned long long* phase;  // optional: wall-clock ticks of sweep/exchange/ICM (workgroup 0)
    uint32_t* fault;         // host-mapped fault words: [0] a cooperative wait gave up, [1] histogram
                             // samples outside their table
    long long* obs_i;        // per slot: samples, sum E, sum up, sum |2 up - n| (k_observe)
    double* obs_d;           // per slot: sum E^2, sum up^2
    uint64_t seed;
    uint32_t tag_base;       // 0 or kTagPilot
    int32_t sweeps;
    int32_t spin_stride;     // LDS bytes per wave for spin bytes / ICM scratch
    int32_t thr_stride;      // LDS words per wave for the threshold row
    int32_t pilot;
    int32_t no_exchange;
    int32_t hist_on;
    int64_t hist_start;
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wave-wide sum and maximum with DPP row shifts and row broadcasts (a
// prefix scan whose last lane holds the total; VALU only, no LDS crossbar).
// Call with all 64 lanes active.
__device__ __forceinline__ int wave_prefix_i32(int v) {  // inclusive: lane i gets v_0 + ... + v_i
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xe, false);  // row_shr:4, banks 1-3
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xc, false);  // row_shr:8, banks 2-3
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    return v;
}
__device__ __forceinline__ int wave_sum_i32(int v) { return __builtin_amdgcn_readlane(wave_prefix_i32(v), 63); }
// Sum over aligned groups of T lanes (T = 1, 2 or 4) of a quad with DPP quad
// permutations; every lane of a group gets the group's sum.  The lanes of a
// group must be active together.
template <int T>
__device__ __forceinline__ int group_sum_i32(int v) {
    if (T >= 2) v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
    if (T >= 4) v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
    return v;
}
__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
// (the identity INT_MIN fills lanes without a source)
__device__ __forceinline__ int wave_max_i32(int v) {
    const int lo = (int)0x80000000;
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x114, 0xf, 0xe, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x118, 0xf, 0xc, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// ---------------------------------------------------------------------------
// Adjacency access.  info[p] = (row start in 16-byte units) << 9 | degree,
// relative to base (LDS for the fused kernel, global memory otherwise).  The
// field of a site reads its row with 128-bit loads, so 64 lanes reading their
// own rows touch consecutive banks instead of one bank per row stride.
// ---------------------------------------------------------------------------
template <int Fmt>
struct Csr {
    const uint32_t* info;
    const uint8_t* base;
    int neg = 1;    // 0: no coupling is negative (unit format: every sign bit is 0, agree = up neighbours)
    int kbit = 15;  // unit format: a record's sign bit (below it the position)
    int tab = 0;    // 1: sp holds the sign table (negated bytes at 2^kbit): a record addresses b ^ sign
    uint32_t sent = 0;  // unit format: the sentinel position 32 * nw (its byte is 0, its sign bit 0)
    __device__ __forceinline__ uint32_t pos_mask() const { return (1u << kbit) - 1u; }
    static constexpr int kPerChunk = Fmt == kFmtUnit ? 8 : 4;
    __device__ __forceinline__ int deg(int p) const { return (int)(info[p] & (uint32_t)kMaxDegree); }
    __device__ __forceinline__ const uint8_t* row(int p) const { return base + (size_t)(info[p] >> 9) * 16u; }
    // Records of 16-byte chunk ch of a row (padding records read position 0).
    __device__ __forceinline__ void chunk(const uint8_t* row, int ch, int* c, int* w) const {
        const uint4 v = reinterpret_cast<const uint4*>(row)[ch];
        const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
        if (Fmt == kFmtUnit) {
#pragma unroll
            for (int h = 0; h < 8; ++h) {
                const uint32_t e = (wv[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
                c[h] = (int)(e & pos_mask());
                w[h] = ((e >> kbit) & 1u) ? -1 : 1;
            }
        } else {
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                c[h] = (int)(wv[h] & 0xFFFFu);
                w[h] = (int32_t)wv[h] >> 16;
            }
        }
    }
    // Neighbour bytes of up to R chunks of a row at once: chunk c0 + k * step
    // (k < R) when below nch, else nothing (masked lanes issue no LDS
    // request).  NEG: records carry a sign bit (agree = byte ^ sign), else a
    // record is the neighbour's position itself.  Returns the number of
    // agreeing neighbours; padding records read the zero sentinel byte with
    // sign 0 and add nothing.
    template <int R, bool NEG>
    __device__ __forceinline__ int agree_chunks(const uint4* r, int c0, int step, int nch, const uint8_t* sp) const {
        uint4 v[R];
#pragma unroll
        for (int k = 0; k < R; ++k) v[k] = c0 + k * step < nch ? r[c0 + k * step] : make_uint4(0u, 0u, 0u, 0u);
        uint32_t b[R][8];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
            if (c0 + k * step < nch) {
#pragma unroll
                for (int h = 0; h < 8; ++h)
                    b[k][h] = NEG ? sp[(wv[h >> 1] >> (16 * (h & 1))) & pos_mask()]
                                  : sp[h & 1 ? wv[h >> 1] >> 16 : wv[h >> 1] & 0xFFFFu];
            } else {
#pragma unroll
                for (int h = 0; h < 8; ++h) b[k][h] = 0u;
            }
        }
        int agree = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
            for (int h = 0; h < 8; ++h)
                agree += NEG ? (int)(b[k][h] ^ ((wv[h >> 1] >> (16 * (h & 1) + kbit)) & 1u)) : (int)b[k][h];
        }
        return agree;
    }
    // agree_chunks without masks: chunk c0 + k * step of the row when below
    // nch (k = 0 always is), else a chunk of sentinel records, whose bytes
    // are 0 and add nothing.  Every gather is issued for every lane (the
    // sentinel reads are one broadcast address) and nothing branches, so the
    // byte loads of all R chunks are in flight together.
    template <int R, bool NEG>
    __device__ __forceinline__ int agree_chunks_s(const uint4* r, int c0, int step, int nch, const uint8_t* sp) const {
        const uint32_t s2 = sent * 0x10001u;
        uint4 v[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (k == 0) {
                v[k] = r[c0];
            } else {
                const uint4 t = r[min(c0 + k * step, nch - 1)];  // a chunk of the row either way
                const bool in = c0 + k * step < nch;
                v[k] = make_uint4(in ? t.x : s2, in ? t.y : s2, in ? t.z : s2, in ? t.w : s2);
            }
        }
        int agree = 0;
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
            uint32_t b[8];
#pragma unroll
            for (int h = 0; h < 8; ++h)
                b[h] = NEG ? sp[(wv[h >> 1] >> (16 * (h & 1))) & pos_mask()]
                           : sp[h & 1 ? wv[h >> 1] >> 16 : wv[h >> 1] & 0xFFFFu];
#pragma unroll
            for (int h = 0; h < 8; ++h)
                agree += NEG ? (int)(b[h] ^ ((wv[h >> 1] >> (16 * (h & 1) + kbit)) & 1u)) : (int)b[h];
        }
        return agree;
    }
    // sum_j w_pj s_j with s = 2b - 1, from byte spins
    __device__ __forceinline__ int field(int p, const uint8_t* sp) const {
        const uint32_t ri = info[p];
        const int d = (int)(ri & (uint32_t)kMaxDegree);
        const uint4* r = reinterpret_cast<const uint4*>(base + (size_t)(ri >> 9) * 16u);
        if (Fmt == kFmtUnit) {
            // Rows of up to four chunks (32 records) load every chunk and
            // issue every gather in one round, longer rows four chunks per
            // round; a lane's chunks beyond its row are masked, so the
            // gathers cost the row's chunks, not the wave's longest row.
            const int nch = (d + 7) >> 3;
            int agree = 0;  // neighbours with s_j * sign(w_pj) = +1
            if (neg) {
                for (int c0 = 0; c0 < nch; c0 += 4) agree += agree_chunks<4, true>(r, c0, 1, nch, sp);
            } else {
                for (int c0 = 0; c0 < nch; c0 += 4) agree += agree_chunks<4, false>(r, c0, 1, nch, sp);
            }
            return 2 * agree - d;
        }
        int f = 0;
        for (int c = 0; 4 * c < d; ++c) {
            const uint4 v = r[c];
            const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int w = (int32_t)wv[h] >> 16;  // 0 on padding records
                f += sp[wv[h] & 0xFFFFu] ? w : -w;
            }
        }
        return f;
    }
    // Part of the local field of p from the row's chunks sub, sub + T, ...
    // (lane sub of T lanes sharing the position): unit format the number of
    // agreeing neighbours (field = 2 * total - deg), wide format the partial
    // sum of w_pj s_j.  Padding records add nothing to either.
    template <int T>
    __device__ __forceinline__ int field_part(int p, const uint8_t* sp, int sub, int& deg) const {
        const uint32_t ri = info[p];
        const int d = (int)(ri & (uint32_t)kMaxDegree);
        deg = d;
        const uint4* r = reinterpret_cast<const uint4*>(base + (size_t)(ri >> 9) * 16u);
        int acc = 0;
        if (Fmt == kFmtUnit) {
            // 4 / T chunks per lane per round: one round for rows of up to 32 records
            constexpr int R = T >= 4 ? 1 : 4 / T;
            const int nch = (d + 7) >> 3;
            // (the loop bound keeps chunk c0 in the row; the round's later
            // chunks are sentinel-filled past it)
            if (neg && !tab) {
                for (int c0 = sub; c0 < nch; c0 += T * R) acc += agree_chunks_s<R, true>(r, c0, T, nch, sp);
            } else {
                for (int c0 = sub; c0 < nch; c0 += T * R) acc += agree_chunks_s<R, false>(r, c0, T, nch, sp);
            }
            return acc;
        }
        for (int c = sub; kPerChunk * c < d; c += T) {
            const uint4 v = r[c];
            const uint32_t wv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const int w = (int32_t)wv[h] >> 16;  // 0 on padding records
                acc += sp[wv[h] & 0xFFFFu] ? w : -w;
            }
        }
        return acc;
    }
    // Local fields of the four consecutive positions p0..p0+3 (positions
    // outside [cs, ce) get 0).  The row words, chunks and neighbour bytes of
    // all four are loaded before any is used, so their latencies overlap;
    // padding records read the zero sentinel byte and need no mask.
    __device__ __forceinline__ void group_field(int p0, int cs, int ce, const uint8_t* sp, int f[4]) const {
        // the sign decision hoisted out of the loops: one loop body per case
        // (with the sign table a signed record reads b ^ sign itself)
        if (Fmt == kFmtUnit && (!neg || tab))
            group_field_t<false>(p0, cs, ce, sp, f);
        else
            group_field_t<true>(p0, cs, ce, sp, f);
    }
    // group_field when every row of the instance is one 16-byte chunk (max
    // degree <= kPerChunk): the rows then sit at a fixed stride, row p being
    // chunk p of the instance's adjacency, so the chunk loads need not wait
    // for the row words (which give only the degrees) and there is no chunk
    // loop.  Positions outside [cs, ce) get 0.
    __device__ __forceinline__ void group_field_one(int p0, int cs, int ce, const uint8_t* sp, int f[4]) const {
        if (Fmt == kFmtUnit && (!neg || tab))
            group_field_one_t<false>(p0, cs, ce, sp, f);
        else
            group_field_one_t<true>(p0, cs, ce, sp, f);
    }
    template <bool NEG>
    __device__ __forceinline__ void group_field_one_t(int p0, int cs, int ce, const uint8_t* sp, int f[4]) const {
        // every load is issued for every lane (a valid row either way) and
        // positions outside the class are masked at the end, so nothing branches
        uint4 v[4];
        int d[4], in[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int p = p0 + k;
            in[k] = -(int)(p >= cs && p < ce);
            const int q = in[k] ? p : cs;
            v[k] = reinterpret_cast<const uint4*>(base)[q];
            d[k] = (int)(info[q] & (uint32_t)kMaxDegree);
        }
        if (Fmt == kFmtUnit) {
            uint32_t b[4][8];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
                for (int h = 0; h < 8; ++h)
                    b[k][h] = NEG ? sp[(wv[h >> 1] >> (16 * (h & 1))) & pos_mask()]
                                  : sp[h & 1 ? wv[h >> 1] >> 16 : wv[h >> 1] & 0xFFFFu];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
                int a = 0;
#pragma unroll
                for (int h = 0; h < 8; ++h)
                    a += NEG ? (int)(b[k][h] ^ ((wv[h >> 1] >> (16 * (h & 1) + kbit)) & 1u)) : (int)b[k][h];
                f[k] = (2 * a - d[k]) & in[k];
            }
        } else {
            uint32_t b[4][4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
                for (int h = 0; h < 4; ++h) b[k][h] = sp[wv[h] & 0xFFFFu];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
                int a = 0;
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    const int w = (int32_t)wv[h] >> 16;
                    a += b[k][h] ? w : -w;
                }
                f[k] = a & in[k];
            }
        }
    }
    template <bool NEG>
    __device__ __forceinline__ void group_field_t(int p0, int cs, int ce, const uint8_t* sp, int f[4]) const {
        uint32_t ri[4];
        int d[4], nch[4];
        int mx = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int p = p0 + k;
            const bool in = p >= cs && p < ce;
            const uint32_t v = info[in ? p : cs];
            ri[k] = in ? v : (v & ~(uint32_t)kMaxDegree);  // out of the class: degree 0, a valid row
            d[k] = (int)(ri[k] & (uint32_t)kMaxDegree);
            nch[k] = (d[k] + kPerChunk - 1) / kPerChunk;
            mx = max(mx, nch[k]);
            f[k] = 0;
        }
        for (int ch = 0; ch < mx; ++ch) {
            uint4 v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int cc = min(ch, max(nch[k] - 1, 0));
                v[k] = reinterpret_cast<const uint4*>(base + (size_t)(ri[k] >> 9) * 16u)[cc];
            }
            if (Fmt == kFmtUnit) {
                uint32_t b[4][8];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
                    for (int h = 0; h < 8; ++h)
                        b[k][h] = NEG ? sp[(wv[h >> 1] >> (16 * (h & 1))) & pos_mask()]
                                      : sp[h & 1 ? wv[h >> 1] >> 16 : wv[h >> 1] & 0xFFFFu];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
                    int a = 0;
#pragma unroll
                    for (int h = 0; h < 8; ++h)
                        a += NEG ? (int)(b[k][h] ^ ((wv[h >> 1] >> (16 * (h & 1) + kbit)) & 1u)) : (int)b[k][h];
                    f[k] += ch < nch[k] ? a : 0;
                }
            } else {
                uint32_t b[4][4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
                    for (int h = 0; h < 4; ++h) b[k][h] = sp[wv[h] & 0xFFFFu];
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t wv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
                    int a = 0;
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const int w = (int32_t)wv[h] >> 16;
                        a += b[k][h] ? w : -w;
                    }
                    f[k] += ch < nch[k] ? a : 0;
                }
            }
        }
        if (Fmt == kFmtUnit) {
#pragma unroll
            for (int k = 0; k < 4; ++k) f[k] = 2 * f[k] - d[k];
        }
    }
};

// Bytes <-> bits.  A state word covers 32 consecutive positions; the byte
// buffer has (nw + 1) * 32 entries: the last 32 (the sentinel word the
// padding records read) and the padding bytes past n stay 0.  lane/stride
// let several waves share the work.
__host__ __device__ inline int spin_bytes(int nw) { return 32 * (nw + 1); }
// Copy of count elements from global memory into LDS by threads tid, tid +
// nthr, ...: each thread issues up to 8 loads before its first store, so a
// workgroup's staging costs about one memory round trip per 8 elements a
// thread copies instead of one per element (a load feeding the LDS store
// of the same iteration serialises the loop on memory latency).
// (Every load is unconditional, from an index clamped into the source, so the
// staging registers stay registers; only the stores are conditional.)
template <typename T>
__device__ __forceinline__ void stage_to_lds(T* __restrict__ dst, const T* __restrict__ src, int count, int tid,
                                             int nthr) {
    constexpr int U = 8;
    if (count <= 0) return;
    for (int q0 = tid; q0 < count; q0 += U * nthr) {
        T v[U];
#pragma unroll
        for (int k = 0; k < U; ++k) v[k] = src[min(q0 + k * nthr, count - 1)];
#pragma unroll
        for (int k = 0; k < U; ++k)
            if (q0 + k * nthr < count) dst[q0 + k * nthr] = v[k];
    }
}
// The position padding records name: the first byte of the sentinel word.
__host__ __device__ inline uint32_t sentinel_pos(int nw) {
    return 32u * (uint32_t)nw < 65535u ? 32u * (uint32_t)nw : 65535u;
}
// Sign table of the split sweep: the spin bytes at [0, spin_bytes) and their
// negations at [2^K, 2^K + spin_bytes), K the smallest bit with 2^K >=
// spin_bytes(nw) (so every position, the sentinel included, is below 2^K).
// A unit record pos | sign << K then addresses b_pos ^ sign directly.
__host__ __device__ inline int sign_table_bit(int nw) {
    int k = 2;
    while ((1 << k) < spin_bytes(nw)) ++k;
    return k;
}
__host__ __device__ inline int sign_table_bytes(int nw) { return (1 << sign_table_bit(nw)) + spin_bytes(nw); }
constexpr int kSignTableMax = 16 * 1024;  // per wave; larger instances keep sign bit 15 and decode it

__device__ __forceinline__ void expand_bits(const uint32_t* __restrict__ words, int nw, uint8_t* sp, int lane,
                                            int stride = kWave) {
    for (int w = lane; w <= nw; w += stride) {
        const uint32_t v = w < nw ? words[w] : 0u;
        uint32_t* dst = reinterpret_cast<uint32_t*>(sp + 32 * w);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t x = (v >> (4 * q)) & 0xFu;
            dst[q] = (x | (x << 7) | (x << 14) | (x << 21)) & 0x01010101u;
        }
    }
}
__device__ __forceinline__ void pack_bits(const uint8_t* sp, int nw, uint32_t* __restrict__ words, int lane,
                                          int stride = kWave) {
    for (int w = lane; w < nw; w += stride) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(sp + 32 * w);
        uint32_t v = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t y = src[q];
            const uint32_t x = (y & 1u) | ((y >> 7) & 2u) | ((y >> 14) & 4u) | ((y >> 21) & 8u);
            v |= x << (4 * q);
        }
        words[w] = v;
    }
}

// ---------------------------------------------------------------------------
// Metropolis sweep of one replica in colour-class order (src/replica.cc:5-38
// restated for class-parallel updates).  Each lane takes aligned groups of 4
// positions and draws one Philox block per group (word k for position 4g+k);
// dE <= 0 is accepted, dE > 0 iff draw < thr[dE/2 - 1].  After every class
// that accepted a flip the best energy is updated (replica.cc:23-27 at class
// granularity).
// ---------------------------------------------------------------------------
struct SweepState {
    int32_t E, bestE, best_sweep;
    unsigned long long flips;
};

// Metropolis updates of the positions 4g..4g+3 that lie in the class
// [cs, ce): one Philox block serves the four (word k for position 4g+k);
// the four spin bytes are read and written as one 32-bit word (no other lane
// owns any byte of it during the class).
template <int Fmt, bool ONE = false>
__device__ __forceinline__ void group_update(const Csr<Fmt>& csr, int g, int cs, int ce,
                                             const int32_t* __restrict__ hf, bool has_field, uint32_t c3, uint32_t k0,
                                             uint32_t k1, int r, uint32_t gsweep, uint8_t* sp,
                                             const uint32_t* thr_row, int& dsum, int& nacc,
                                             const uint32_t* pre = nullptr, uint8_t* spn = nullptr) {
    // spn (optional): the sign table's negated bytes, kept equal to sp ^ 1
    // pre (optional): this sweep's random words, pre[p] = word p&3 of block p>>2
    const int p0 = 4 * g;
    int f[4];
    if (ONE)
        csr.group_field_one(p0, cs, ce, sp, f);
    else
        csr.group_field(p0, cs, ce, sp, f);
    const uint32_t own = *reinterpret_cast<const uint32_t*>(sp + p0);
    int dE[4];
    bool any_up = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = p0 + k;
        const bool in = p >= cs && p < ce;
        int fk = f[k];
        if (ONE) {  // has_field is uniform over the wave: one scalar branch, unconditional loads
            if (has_field) fk += hf[in ? p : cs] & -(int)in;
        } else {
            fk += (has_field && in) ? hf[p] : 0;
        }
        const int m = -(int)((own >> (8 * k)) & 1u);  // -1 for a spin set: dE = -2 f, else 2 f
        dE[k] = in ? 2 * ((fk ^ m) - m) : 0;
        any_up |= dE[k] > 0;
    }
    Philox4 rnd = {{0u, 0u, 0u, 0u}};
    if (any_up) {
        if (pre) {
            const uint4 q = *reinterpret_cast<const uint4*>(pre + p0);
            rnd.v[0] = q.x;
            rnd.v[1] = q.y;
            rnd.v[2] = q.z;
            rnd.v[3] = q.w;
        } else {
            rnd = philox4x32_10((uint32_t)g, (uint32_t)r, gsweep, c3, k0, k1);
        }
    }
    uint32_t flip = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = p0 + k;
        if (ONE) {
            // threshold read for every position (entry 0 when dE <= 0, where
            // it is not used; the row holds at least one word), no branches
            // dE is even: entry dE/2 - 1 sits at byte 2 dE - 4
            const uint32_t t =
                *reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint8_t*>(thr_row) + 2 * max(dE[k], 2) - 4);
            const uint32_t acc = (uint32_t)(p >= cs && p < ce && (dE[k] <= 0 || rnd.v[k] < t));
            flip |= acc << (8 * k);
            dsum += dE[k] & -(int)acc;
            nacc += (int)acc;
        } else if (p >= cs && p < ce && (dE[k] <= 0 || rnd.v[k] < thr_row[(dE[k] >> 1) - 1])) {
            flip |= 1u << (8 * k);
            dsum += dE[k];
            ++nacc;
        }
    }
    if (flip) {
        *reinterpret_cast<uint32_t*>(sp + p0) = own ^ flip;
        if (spn) *reinterpret_cast<uint32_t*>(spn + p0) = own ^ flip ^ 0x01010101u;
    }
}

template <int Fmt, bool ONE = false>
__device__ __forceinline__ void sweep_replica(const Csr<Fmt>& csr, const int32_t* __restrict__ cls, int ncls,
                                              const int32_t* __restrict__ hf, bool has_field, int nw, uint32_t c3,
                                              uint64_t seed, int r, uint32_t gsweep, uint8_t* sp,
                                              const uint32_t* thr_row, uint32_t* best_words, SweepState& st,
                                              int lane) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int c = 0; c < ncls; ++c) {
        const int cs = cls[c], ce = cls[c + 1];
        int dsum = 0, nacc = 0;
        uint8_t* spn = csr.tab ? sp + (1 << csr.kbit) : nullptr;
        for (int g = (cs >> 2) + lane; 4 * g < ce; g += kWave) {
            group_update<Fmt, ONE>(csr, g, cs, ce, hf, has_field, c3, k0, k1, r, gsweep, sp, thr_row, dsum, nacc,
                                   nullptr, spn);
        }
        wave_sync();
        dsum = wave_sum_i32(dsum);
        nacc = wave_sum_i32(nacc);
        st.E += dsum;
        st.flips += (unsigned long long)nacc;
        if (nacc > 0 && st.E < st.bestE) {
            st.bestE = st.E;
            st.best_sweep = (int32_t)gsweep;
            pack_bits(sp, nw, best_words, lane);
        }
    }
}

// ---------------------------------------------------------------------------
// Replica-sliced sweep (InstDesc::rs: unit records, no field, degree <= 63).
// The W waves of a workgroup sweep 32 replicas of an instance together.  LDS
// word X[p] holds position p's spins in the 32 replicas (bit j: replica
// r0 + j), so one gather of a neighbour's word serves all 32, and a
// position's agreeing-neighbour counts are kept bit-sliced: six planes P[i]
// whose bit j is bit i of replica j's count.  The decisions are
// sweep_replica's: thread t of the 64 W takes the groups of 4 positions
// (cs >> 2) + t, + 64 W, ... of a class, and replica r's update of position
// 4g + k draws word k of the Philox block (g, r, sweep) and tests it against
// the threshold row of r's temperature, so spins, energies, flip counts and
// best states equal the per-replica sweep's.
// ---------------------------------------------------------------------------
// 32 x 32 bit transpose across the 32 lanes of a half-wave (all active):
// lane i's result has bit j = bit i of lane j's v.  Five rounds of the
// recursive block transpose (Hacker's Delight, section 7-3, low bit first):
// in round J, lanes k and k ^ J exchange the high J bits of each 2J-bit
// field of the lower lane with the low J bits of the upper lane.
__device__ __forceinline__ uint32_t transpose_lanes32(uint32_t v, int l32) {
#pragma unroll
    for (int J = 16; J >= 1; J >>= 1) {
        const uint32_t m = J == 16 ? 0x0000FFFFu
                           : J == 8 ? 0x00FF00FFu
                           : J == 4 ? 0x0F0F0F0Fu
                           : J == 2 ? 0x33333333u
                                    : 0x55555555u;
        const uint32_t u = (uint32_t)__shfl_xor((int)v, J, 32);
        const bool hi = (l32 & J) != 0;
        const uint32_t t = hi ? (((u >> J) ^ v) & m) : (((v >> J) ^ u) & m);
        v ^= hi ? t : (t << J);
    }
    return v;
}
// a + b + c = s + 2 co, bitwise (co: the majority, a where a == b, else c)
__device__ __forceinline__ void full_add(uint32_t a, uint32_t b, uint32_t c, uint32_t& s, uint32_t& co) {
    const uint32_t u = a ^ b;
    s = u ^ c;
    co = (u & c) | (~u & a);
}
// P += A[0] + ... + A[7], bit-sliced (per bit position j: six-bit counters).
// A carry-save tree reduces the eight words to a four-bit sum, which a
// ripple carry adds to the planes.  Counts stay below 64 (degree <= 63).
__device__ __forceinline__ void add8_planes(uint32_t (&P)[6], const uint32_t (&A)[8]) {
    uint32_t s1, c1, s2, c2, s0, d1, t1, e1;
    full_add(A[0], A[1], A[2], s1, c1);
    full_add(A[3], A[4], A[5], s2, c2);
    const uint32_t s3 = A[6] ^ A[7], c3 = A[6] & A[7];
    full_add(s1, s2, s3, s0, d1);  // weight 1: s0; weight 2: c1, c2, c3, d1
    full_add(c1, c2, c3, t1, e1);  // weight 2: t1, d1; weight 4: e1
    const uint32_t b1 = t1 ^ d1, e2 = t1 & d1;
    const uint32_t b2 = e1 ^ e2, b3 = e1 & e2;  // the sum: b3 b2 b1 s0
    uint32_t cy = P[0] & s0, n;
    P[0] ^= s0;
    full_add(P[1], b1, cy, n, cy);
    P[1] = n;
    full_add(P[2], b2, cy, n, cy);
    P[2] = n;
    full_add(P[3], b3, cy, n, cy);
    P[3] = n;
    n = P[4] ^ cy;
    cy &= P[4];
    P[4] = n;
    P[5] ^= cy;
}

// Shared state of the W waves sweeping one block of replicas: per class the
// waves' partial sums per replica and the mask of replicas with a new best.
struct SlicedShared {
    int32_t dsum[8][32];
    int32_t nacc[8][32];
    uint32_t best_mask;
};

constexpr int kSlicedShared = (int)((sizeof(SlicedShared) + 15) / 16 * 16);
// LDS bytes of k_sweep_rs for an instance of nw state words and threshold
// rows of fmax entries
__host__ __device__ inline int64_t sliced_lds_bytes(int nw, int fmax) {
    return kSlicedShared + 4ll * 32 * (nw + 1) + 4ll * 32 * (fmax > 1 ? fmax : 1);
}

// Replica j's count from six bit planes (P[i] bit j = bit i of the count).
__device__ __forceinline__ int plane_count(const uint32_t (&P)[6], int j) {
    int c = (int)__builtin_amdgcn_ubfe(P[0], (uint32_t)j, 1u);
#pragma unroll
    for (int i = 1; i < 6; ++i) c |= (int)__builtin_amdgcn_ubfe(P[i], (uint32_t)j, 1u) << i;
    return c;
}

// Sweeps of replicas r0 .. r0 + nb - 1 (nb <= 32) by the W waves of a
// workgroup (thread tid of 64 W), the replicas' position words in X (32 (nw
// + 1) words: the state's blocks, then the zero block of the sentinel
// position), threshold rows in T (row j: replica r0 + j).  The groups of a
// class are dealt to the 64 W lanes; wave 0's lane j < nb keeps replica
// r0 + j's energy, best energy, best sweep and flip count in st.
template <int W>
__device__ __forceinline__ void sweep_sliced(const InstDesc& d, const uint32_t* __restrict__ info,
                                             const uint4* __restrict__ adj, const int32_t* __restrict__ cls,
                                             uint32_t c3, uint32_t k0, uint32_t k1, int r0, int nb,
                                             uint32_t gsweep, uint32_t* X, const uint32_t* T, SlicedShared& sh,
                                             uint32_t* __restrict__ best_words, SweepState& st, int tid) {
    const int kb = d.kbit, nw = d.nw, wave = tid / kWave, lane = tid % kWave;
    const uint32_t pmask = (1u << kb) - 1u;
    const uint32_t sent2 = sentinel_pos(nw) * 0x10001u;  // two sentinel records (position 32 nw, sign 0)
    for (int c = 0; c < d.ncls; ++c) {
        const int cs = cls[c], ce = cls[c + 1];
        int dacc = 0, nacc = 0;  // lane j: replica r0 + j's sums over this wave's groups of the class
        for (int g0 = cs >> 2; 4 * g0 < ce; g0 += W * kWave) {  // uniform bound: every lane joins the reductions
            const int g = g0 + tid, p0 = 4 * g;
            uint32_t P[4][6], own[4];
            int deg[4], nch[4], inm[4], mx = 0;  // inm: -1 for a position of the class, else 0
            const uint4* row[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int p = p0 + k;
                const bool in = p >= cs && p < ce;
                inm[k] = -(int)in;
                const uint32_t ri = info[in ? p : cs];  // a valid entry either way
                deg[k] = (int)(ri & (uint32_t)kMaxDegree) & inm[k];
                row[k] = adj + (ri >> 9);
                nch[k] = (deg[k] + 7) >> 3;
                mx = max(mx, nch[k]);
                own[k] = X[in ? p : cs];  // (lanes past the class end read a valid word)
#pragma unroll
                for (int i = 0; i < 6; ++i) P[k][i] = 0u;
            }
            // two chunks of every row per round, all loads before any gather;
            // chunks past a row's end are replaced by sentinel records
            for (int c0 = 0; c0 < mx; c0 += 2) {
                uint4 q[4][2];
#pragma unroll
                for (int k = 0; k < 4; ++k)
#pragma unroll
                    for (int h = 0; h < 2; ++h) q[k][h] = row[k][min(c0 + h, max(nch[k] - 1, 0))];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const bool live = c0 + h < nch[k];
                        const uint32_t wv[4] = {live ? q[k][h].x : sent2, live ? q[k][h].y : sent2,
                                                live ? q[k][h].z : sent2, live ? q[k][h].w : sent2};
                        uint32_t A[8];
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            // agreeing neighbours: spin bits, inverted for a negative coupling
                            // (sign 0 or -1 by a signed one-bit extract); sentinel records read
                            // the zero block with sign 0
                            const uint32_t w32 = wv[e >> 1];
                            const uint32_t pos = (e & 1) ? __builtin_amdgcn_ubfe(w32, 16u, (uint32_t)kb) : (w32 & pmask);
                            const uint32_t sgn =
                                (uint32_t)__builtin_amdgcn_sbfe((int)w32, (e & 1) ? 16u + (uint32_t)kb : (uint32_t)kb, 1u);
                            A[e] = X[pos] ^ sgn;
                        }
                        add8_planes(P[k], A);
                    }
                }
            }
            uint32_t F[4] = {0u, 0u, 0u, 0u};
            for (int j = 0; j < nb; ++j) {
                int dE[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int f = 2 * plane_count(P[k], j) - deg[k];
                    const int m = -(int)__builtin_amdgcn_ubfe(own[k], (uint32_t)j, 1u);  // -1: spin set, dE = -2 f
                    dE[k] = (2 * ((f ^ m) - m)) & inm[k];
                }
                const Philox4 rnd = philox4x32_10((uint32_t)g, (uint32_t)(r0 + j), gsweep, c3, k0, k1);
                const uint32_t* tr = T + j * d.fmax;
                int v = 0;  // accepted dE + 65536 per acceptance (|sum dE| < 32768 over the wave)
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint32_t t = tr[max(dE[k] >> 1, 1) - 1];  // entry 0 where dE <= 0 (not used)
                    // no short-circuit: selects, not branches
                    const uint32_t acc =
                        (uint32_t)inm[k] & ((uint32_t)(dE[k] <= 0) | (uint32_t)(rnd.v[k] < t));
                    F[k] |= acc << j;
                    v += (dE[k] + 65536) & -(int)acc;
                }
                v = wave_sum_i32(v);
                const int na = (v + 32768) >> 16;
                dacc += lane == j ? v - na * 65536 : 0;
                nacc += lane == j ? na : 0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (F[k]) X[p0 + k] = own[k] ^ F[k];
        }
        if (lane < 32) {
            sh.dsum[wave][lane] = dacc;
            sh.nacc[wave][lane] = nacc;
        }
        __syncthreads();
        if (wave == 0) {
            int ds = 0, na = 0;
            if (lane < nb) {
#pragma unroll
                for (int x = 0; x < W; ++x) {
                    ds += sh.dsum[x][lane];
                    na += sh.nacc[x][lane];
                }
            }
            st.E += ds;
            st.flips += (unsigned long long)na;
            const bool better = lane < nb && na > 0 && st.E < st.bestE;
            if (better) {
                st.bestE = st.E;
                st.best_sweep = (int32_t)gsweep;
            }
            const unsigned long long bm = __ballot(better);
            if (lane == 0) sh.best_mask = (uint32_t)bm;
        }
        __syncthreads();
        const uint32_t bm = sh.best_mask;
        if (bm) {
            // half-wave hw packs state word w = hw, hw + 2 W, ...: after the
            // transpose its lane j holds replica r0 + j's word w
            const int l32 = tid & 31;
            for (int w = tid >> 5; w < nw; w += 2 * W) {
                const uint32_t v = transpose_lanes32(X[32 * w + l32], l32);
                if ((bm >> l32) & 1u) best_words[(int64_t)l32 * nw + w] = v;
            }
            __syncthreads();  // the next class writes X
        }
    }
}

// ---------------------------------------------------------------------------
// Multispin sweep of a periodic lattice (InstDesc::lsw_dim = D > 0): one wave
// per replica, the state as position-order bits in LDS, the same decisions
// as sweep_replica (Replica::sweep, src/replica.cc:5-38, in colour order).
//
// The two colour classes are the site parities, each listing its sites in
// index order: the class-c sites of lattice row r = y + L z are the H = L/2
// consecutive positions class_start[c] + r H + k (site x = 2k + par, par =
// (c + y + z) & 1).  So the neighbours of class position q in the other class
// sit at q + delta for a few fixed offsets: 0 and +1 or -(H-1) (+x), 0 and -1
// or +(H-1) (-x), by par and by k = H-1 or 0; +-H or -+(L-1) H along y and
// +-L H or -+(L-1) L H along z, the second value where the axis wraps.  A
// colour class is swept in two phases:
//  A  one lane per window of 32 class positions: the neighbours' bits are 32-
//     bit fields of the other class's bits at those offsets, chosen per bit by
//     the window's shape masks; the bond bits B = ~(s ^ s' ^ neg), 1 where
//     w s s' = +1 (dE = 2 deg - 4 count(B)), are counted in three bit planes;
//     sites with count >= D (dE <= 0) flip at once, and the others' uphill
//     level j = count (dE = 4 (D - j)) goes to LDS as D bit masks;
//  B  one lane per Philox block of 4 positions (block p >> 2, word p & 3: the
//     word every sweep draws for position p): an uphill site flips when its
//     word is below the threshold of its level.
// Flips are XORs into the state words (windows and blocks share words).
// ---------------------------------------------------------------------------
// 32 bits of S from bit p.  A field whose bits are used starts within 31
// positions of the state (its used bits lie inside it), so S has a zero word
// before it and two after; a field none of whose bits is used may start
// anywhere, and p is clamped into [-32, n] for it.
__device__ __forceinline__ uint32_t lsw_bits(const uint32_t* S, int p, int n) {
    p = min(max(p, -32), n);
    const int w0 = p >> 5;  // -1 for p in [-32, -1]
    return (uint32_t)((((uint64_t)S[w0 + 1] << 32) | S[w0]) >> (p & 31));
}

__device__ __forceinline__ void lsw_xor(uint32_t* S, int p, uint32_t v) {
    if (!v) return;
    const uint64_t x = (uint64_t)v << (p & 31);
    const int w0 = p >> 5;
    if ((uint32_t)x) atomicXor(S + w0, (uint32_t)x);
    if ((uint32_t)(x >> 32)) atomicXor(S + w0 + 1, (uint32_t)(x >> 32));
}

__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c) { return (a & b) | (c & (a | b)); }

// LDS words per wave: the state (a zero word, nw words, two zero words) and
// D level masks of one class, each with a zero word before and after.
__host__ __device__ inline int lsw_state_words(int nw) { return (nw + 3 + 3) & ~3; }
__host__ __device__ inline int lsw_zwords(int n) { return ((n / 2 + 31) >> 5) + 2; }
__host__ __device__ inline int lsw_lds_bytes(int nw, int n, int D) {
    return 4 * (lsw_state_words(nw) + D * lsw_zwords(n));
}

// Phase A of class c (class starts cs, other class cn): tab holds per
// window 16 words, the edge signs (+x, -x, +y, -y, +z, -z, 0, 0) and the
// shape masks (par, k = 0, k = H-1, y = 0, y = L-1, z = 0, z = L-1, 0).
template <int D>
__device__ __forceinline__ void lsw_phase_a(const uint32_t* __restrict__ tab, int L, int n, int c, int cs, int cn,
                                            uint32_t* S, uint32_t* Z, int zw, int lane, int& dsum, int& nacc) {
    const int H = L >> 1, hc = n >> 1, nwin = (hc + 31) >> 5;
    const int dYw = (L - 1) * H, dZ = L * H, dZw = (L - 1) * L * H;
    for (int w = lane; w < nwin; w += kWave) {
        const int q0 = 32 * w;
        const uint32_t vmask = hc - q0 >= 32 ? ~0u : (1u << (hc - q0)) - 1u;
        const uint4* g = reinterpret_cast<const uint4*>(tab + ((size_t)c * nwin + w) * kLswGroup);
        const uint4 s0 = g[0], s1 = g[1], m0 = g[2], m1 = g[3];
        const int pn = cn + q0;
        const uint32_t own = lsw_bits(S, cs + q0, n);
        const uint32_t f0 = lsw_bits(S, pn, n);
        const uint32_t par = m0.x, k0 = m0.y, kh = m0.z;
        uint32_t N[2 * D];
        N[0] = (par & ((kh & lsw_bits(S, pn - (H - 1), n)) | (~kh & lsw_bits(S, pn + 1, n)))) | (~par & f0);
        N[1] = (par & f0) | (~par & ((k0 & lsw_bits(S, pn + (H - 1), n)) | (~k0 & lsw_bits(S, pn - 1, n))));
        N[2] = (m1.x & lsw_bits(S, pn - dYw, n)) | (~m1.x & lsw_bits(S, pn + H, n));
        N[3] = (m0.w & lsw_bits(S, pn + dYw, n)) | (~m0.w & lsw_bits(S, pn - H, n));
        if (D == 3) {
            N[4] = (m1.z & lsw_bits(S, pn - dZw, n)) | (~m1.z & lsw_bits(S, pn + dZ, n));
            N[5] = (m1.y & lsw_bits(S, pn + dZw, n)) | (~m1.y & lsw_bits(S, pn - dZ, n));
        }
        const uint32_t sg[6] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y};
        uint32_t B[2 * D];
#pragma unroll
        for (int j = 0; j < 2 * D; ++j) B[j] = ~(own ^ N[j] ^ sg[j]);
        // count(B) = c0 + 2 c1 + 4 c2
        uint32_t c0, c1, c2;
        {
            const uint32_t s1a = B[0] ^ B[1] ^ B[2], k1 = maj3(B[0], B[1], B[2]);
            if (D == 3) {
                const uint32_t s2a = B[3] ^ B[4] ^ B[5], k2 = maj3(B[3], B[4], B[5]);
                c0 = s1a ^ s2a;
                const uint32_t k3 = s1a & s2a;
                c1 = k1 ^ k2 ^ k3;
                c2 = maj3(k1, k2, k3);
            } else {
                c0 = s1a ^ B[3];
                const uint32_t k2 = s1a & B[3];
                c1 = k1 ^ k2;
                c2 = k1 & k2;
            }
        }
        const uint32_t dn = (D == 3 ? (c2 | (c1 & c0)) : (c1 | c2)) & vmask;  // count >= D: dE <= 0
        const int na = __popc(dn);
        nacc += na;
        dsum += 4 * D * na - 4 * (__popc(dn & c0) + 2 * __popc(dn & c1) + 4 * __popc(dn & c2));
        const uint32_t nz = ~(c1 | c2);
        Z[1 + w] = nz & ~c0 & vmask;       // count 0
        Z[zw + 1 + w] = nz & c0 & vmask;   // count 1
        if (D == 3) Z[2 * zw + 1 + w] = c1 & ~c2 & ~c0 & vmask;  // count 2
        lsw_xor(S, cs + q0, dn);
    }
}

// Phase B of class c: lane per Philox block g of the class's positions.
template <int D>
__device__ __forceinline__ void lsw_phase_b(int n, int cs, const uint32_t* Z, int zw, const uint32_t* T, uint32_t c3,
                                            uint32_t k0, uint32_t k1, int r, uint32_t gsweep, uint32_t* S, int lane,
                                            int& dsum, int& nacc) {
    const int hc = n >> 1;
    const int gs = cs >> 2, ge = (cs + hc + 3) >> 2;
    for (int g = gs + lane; g < ge; g += kWave) {
        // level bits of the block's positions (class-relative bit 32 + q of the
        // padded masks; the pads are zero, so positions outside the class have none)
        const int b = 32 + 4 * g - cs, bw = b >> 5, bs = b & 31;
        uint32_t z[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            const uint32_t* zj = Z + j * zw;
            z[j] = (uint32_t)((((uint64_t)zj[bw + 1] << 32) | zj[bw]) >> bs) & 0xFu;
        }
        const Philox4 x = philox4x32_10((uint32_t)g, (uint32_t)r, gsweep, c3, k0, k1);
        uint32_t acc = 0;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            uint32_t cj = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) cj |= (x.v[i] < T[j] ? 1u : 0u) << i;
            acc |= cj & z[j];
        }
        if (acc) {
            const int na = __popc(acc);
            nacc += na;
            dsum += 4 * D * na - 4 * __popc(acc & z[1]) - (D == 3 ? 8 * __popc(acc & z[D - 1]) : 0);
            atomicXor(S + (g >> 3), acc << (4 * (g & 7)));
        }
    }
}

template <int D>
__device__ __forceinline__ void sweep_lattice(const uint32_t* __restrict__ tab, int L, int n,
                                              const int32_t* __restrict__ cls, int nw, uint32_t c3, uint64_t seed,
                                              int r, uint32_t gsweep, uint32_t* S, uint32_t* Z,
                                              const uint32_t* thr_row, uint32_t* best_words, SweepState& st,
                                              int lane) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const int zw = lsw_zwords(n);
    // thresholds of the uphill levels: count j gives dE = 4 (D - j), entry 2 (D - j) - 1
    uint32_t T[D];
#pragma unroll
    for (int j = 0; j < D; ++j) T[j] = thr_row[2 * (D - j) - 1];
    for (int c = 0; c < 2; ++c) {
        int dsum = 0, nacc = 0;
        lsw_phase_a<D>(tab, L, n, c, cls[c], cls[1 - c], S, Z, zw, lane, dsum, nacc);
        wave_sync();
        lsw_phase_b<D>(n, cls[c], Z, zw, T, c3, k0, k1, r, gsweep, S, lane, dsum, nacc);
        wave_sync();
        dsum = wave_sum_i32(dsum);
        nacc = wave_sum_i32(nacc);
        st.E += dsum;
        st.flips += (unsigned long long)nacc;
        if (nacc > 0 && st.E < st.bestE) {
            st.bestE = st.E;
            st.best_sweep = (int32_t)gsweep;
            for (int w = lane; w < nw; w += kWave) best_words[w] = S[w];
        }
    }
}

// Energy of the byte state: E = sum_p h_p s_p + 1/2 sum_p s_p sum_q w s_q
// (src/problem.cc:121-142).
template <int Fmt>
__device__ __forceinline__ int32_t energy_bytes(const Csr<Fmt>& csr, int n, const int32_t* hf, bool has_field,
                                                const uint8_t* sp, int lane) {
    long long acc2 = 0;
    for (int p = lane; p < n; p += kWave) {
        const int sgn = sp[p] ? 1 : -1;
        acc2 += (long long)sgn * csr.field(p, sp) + (has_field ? 2ll * sgn * hf[p] : 0ll);
    }
    acc2 = wave_sum_i64(acc2);
    return (int32_t)(acc2 / 2);
}

// ---------------------------------------------------------------------------
// (NN)_a exchange of one instance, executed by one wavefront on LDS state.
//
// The reference walks 2(Nt-1) attempts in a fixed order, each reading the
// permutation the earlier ones left (replica_exchange.cc:5-100 driven by
// solver.cc:81-89; gather_samples' loop in pilot mode, cae_estimator.cc:
// 122-139).  An attempt on pair k reads slots k and k+1, so it depends only on
// the latest earlier attempts on pairs k-1, k and k+1 of the same chain.  The
// attempts of one half (one pair order) are distinct pairs, so they are
// grouped into dependency levels and every level runs in parallel across
// lanes; the result is the sequential one bit for bit.
//
// LDS inputs: beta[nrep] (slot betas), E[nrep] (by replica), s2r[nrep];
// outputs: s2r, att/succ (+=).  Scratch (ExchScratch) sized by nrep.
// ---------------------------------------------------------------------------
struct ExchScratch {
    double* logu;    // [nrep]  log-uniform of attempt a
    uint32_t* key;   // [nrep]  sort keys of one half
    int32_t* perm;   // [nrep]  pair order of the current half
    int32_t* tpos;   // [nrep]  attempt index of pair k in this half, -1 if none
    int32_t* level;  // [nrep]
};

// The same exchange with the permutation held in registers, lane k owning
// slot k (nrep <= 64): sort ranks, dependency levels and the level-by-level
// decisions move between neighbouring slots with shuffles instead of LDS
// round trips and wave barriers.  Bit-identical to the LDS form.
//
// Everything but the decisions depends only on (seed, step): exch_prepare
// computes, per half and lane, the attempt index (tpos), its dependency level
// and its log-uniform; exch_apply makes the decisions from them and the
// energies.  A kernel can prepare step s+1 on one wave while another applies
// step s.
struct ExchPre {
    int32_t tpos[2], level[2];
    double lu[2];
    int32_t maxlv[2];  // the highest level of each half (the same in every lane)
};

// The value of lane + 1 (lane_next) or lane - 1 (lane_prev) by DPP wave
// shifts on the VALU (no LDS crossbar round trip); a lane without such a
// neighbour keeps its own value, as with __shfl_down / __shfl_up by 1.  All
// 64 lanes must be active.
__device__ __forceinline__ int lane_next(int v) {
    return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false);  // wave_shl:1
}
__device__ __forceinline__ int lane_prev(int v) {
    return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);  // wave_shr:1
}

// kbuf: at least nt - 1 words of LDS owned by the calling wave, where the
// sort keys of a half are compared.  only_half >= 0 prepares that half
// alone (the other keeps tpos -1): two waves can prepare the two halves at
// once.  with_lu = false leaves lu at 0 (exch_attempt_logu gives the
// values, by attempt, from another wave).
__device__ __forceinline__ ExchPre exch_prepare(int nt, int nrep, bool pilot, uint64_t seed, uint32_t tag_base,
                                                uint32_t inst_id, uint32_t step, int lane, uint32_t* kbuf,
                                                int only_half = -1, bool with_lu = true) {
    const int np = nt - 1;
    const int nhalf = pilot ? 1 : 2;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t c3s = ctr3(kTagShuf | tag_base, inst_id);
    const uint32_t c3e = ctr3(kTagExch | tag_base, inst_id);
    const bool valid = lane < nrep;
    const int kc = valid ? lane % nt : 0;
    const int my_chain = valid ? lane / nt : 0;
    const bool is_pair = valid && kc < np;  // slot k has a right neighbour in its chain
    ExchPre x;
    for (int h = 0; h < 2; ++h) {
        x.tpos[h] = -1;
        x.level[h] = 0;
        x.lu[h] = 0.0;
        x.maxlv[h] = 0;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // unrolled: x is indexed by constants and stays in registers
        if (h >= nhalf || (only_half >= 0 && h != only_half)) continue;
        // rank of pair kc in this half's order: pairs sorted by (key, index)
        const uint32_t key =
            is_pair ? pick(philox4x32_10((uint32_t)kc >> 2, (uint32_t)h, step, c3s, k0, k1), (uint32_t)kc) : 0u;
        // the keys of pairs 0..np-1 (lanes 0..np-1: chain 0's pairs) in LDS,
        // read back at one address for every lane, four reads in flight
        wave_sync();  // an earlier half's reads of kbuf are done
        if (lane < np) kbuf[lane] = key;
        wave_sync();
        int rank = 0;
        int i = 0;
        for (; i + 4 <= np; i += 4) {
            const uint32_t kk[4] = {kbuf[i], kbuf[i + 1], kbuf[i + 2], kbuf[i + 3]};
#pragma unroll
            for (int j = 0; j < 4; ++j) rank += (kk[j] < key) || (kk[j] == key && i + j < kc);
        }
        for (; i < np; ++i) {
            const uint32_t ki = kbuf[i];
            rank += (ki < key) || (ki == key && i < kc);
        }
        const int a = h * np + rank;
        const int chain = pilot ? 0 : (a & 1);
        const int tpos = (is_pair && chain == my_chain) ? a : -1;
        x.tpos[h] = tpos;
        if (with_lu)
            x.lu[h] = tpos >= 0 ? uniform_log(pick(philox4x32_10((uint32_t)a >> 2, 0u, step, c3e, k0, k1), (uint32_t)a))
                                : 0.0;
        // dependency levels: 1 + the highest level of an earlier attempt on k-1 or k+1
        int level = 0;
        while (true) {
            const int tl = lane_prev(tpos), ll = lane_prev(level);
            const int tr = lane_next(tpos), lr = lane_next(level);
            int lv = 0;
            if (tpos >= 0) {
                lv = 1;
                if (kc > 0 && tl >= 0 && tl < tpos) lv = max(lv, ll + 1);
                if (kc + 1 < np && tr >= 0 && tr < tpos) lv = max(lv, lr + 1);
            }
            const bool changed = __ballot(lv != level) != 0ull;
            level = lv;
            if (!changed) break;
        }
        x.level[h] = level;
        x.maxlv[h] = wave_max_i32(level);
    }
    return x;
}

// The log-uniform of every attempt of the step, by attempt index (lane a
// takes attempt a; at most 64 attempts: 2 (nt - 1) with two chains of at
// most 32 slots, nt - 1 in pilot mode with at most 64): the lu values
// exch_prepare would give the lanes whose tpos is a.
__device__ __forceinline__ void exch_attempt_logu(int nt, bool pilot, uint64_t seed, uint32_t tag_base,
                                                  uint32_t inst_id, uint32_t step, int lane, double* out) {
    const int na = (pilot ? 1 : 2) * (nt - 1);
    const uint32_t c3e = ctr3(kTagExch | tag_base, inst_id);
    if (lane < na)
        out[lane] = uniform_log(
            pick(philox4x32_10((uint32_t)lane >> 2, 0u, step, c3e, (uint32_t)seed, (uint32_t)(seed >> 32)),
                 (uint32_t)lane));
}

// r2s (optional): also written, r2s[replica] = slot, from the same registers.
// rep, es, bk: the lane's slot's replica, its energy and the slot's beta
// (0 for lanes at or past nrep), read by the caller -- a caller can read
// rep and bk, and poll es, before the step's energies are all in.
__device__ __forceinline__ unsigned long long exch_apply(int nt, int nrep, bool pilot, const ExchPre& x, int rep,
                                                         int32_t es, double bk, int32_t* s2r, uint32_t* att,
                                                         uint32_t* succ, int lane, int32_t* r2s = nullptr) {
    const int nhalf = pilot ? 1 : 2;
    const bool valid = lane < nrep;
    const int kc = valid ? lane % nt : 0;
    // beta of slot k + 1 (used only at a decision), by DPP on both halves
    const unsigned long long bkb = (unsigned long long)__double_as_longlong(bk);
    const double bk1 = __longlong_as_double(
        (long long)(((unsigned long long)(uint32_t)lane_next((int)(uint32_t)(bkb >> 32)) << 32) |
                    (uint32_t)lane_next((int)(uint32_t)bkb)));
    const double db = bk - bk1;
    uint32_t natt = 0, nsucc = 0;
    unsigned long long nsw = 0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // unrolled: x is indexed by constants and stays in registers
        if (h >= nhalf) break;
        const int level = x.level[h];
        const int maxlv = __builtin_amdgcn_readfirstlane(x.maxlv[h]);
        for (int L = 1; L <= maxlv; ++L) {
            const int rep_r = lane_next(rep);
            const int32_t es_r = lane_next(es);
            int sw = 0;
            if (level == L) {
                const double lb = db * (double)(es - es_r);
                ++natt;
                if (x.lu[h] < lb) {
                    sw = 1;
                    ++nsucc;
                    ++nsw;
                }
            }
            const int rep_l = lane_prev(rep);
            const int32_t es_l = lane_prev(es);
            const int swl = lane_prev(sw);
            if (sw) {
                rep = rep_r;
                es = es_r;
            } else if (kc > 0 && swl) {  // the pair on the left swapped slots k-1 and k
                rep = rep_l;
                es = es_l;
            }
        }
    }
    if (valid) {
        s2r[lane] = rep;
        if (r2s) r2s[rep] = lane;
        att[lane] += natt;
        succ[lane] += nsucc;
    }
    return nsw;
}

__device__ __forceinline__ unsigned long long exchange_wave_reg(int nt, int nrep, bool pilot, uint64_t seed,
                                                                uint32_t tag_base, uint32_t inst_id, uint32_t step,
                                                                const double* beta, const int32_t* E, int32_t* s2r,
                                                                uint32_t* att, uint32_t* succ, uint32_t* kbuf,
                                                                int lane) {
    const ExchPre x = exch_prepare(nt, nrep, pilot, seed, tag_base, inst_id, step, lane, kbuf);
    const bool valid = lane < nrep;
    const int rep = valid ? s2r[lane] : 0;
    return exch_apply(nt, nrep, pilot, x, rep, valid ? E[rep] : 0, valid ? beta[lane] : 0.0, s2r, att, succ, lane);
}

__device__ __forceinline__ unsigned long long exchange_wave(int nt, int nrep, bool pilot, uint64_t seed,
                                                            uint32_t tag_base, uint32_t inst_id, uint32_t step,
                                                            const double* beta, const int32_t* E, int32_t* s2r,
                                                            uint32_t* att, uint32_t* succ, ExchScratch x, int lane) {
    const int np = nt - 1;
    if (np <= 0) return 0ull;
    if (nrep <= kWave) {
        const unsigned long long r = exchange_wave_reg(nt, nrep, pilot, seed, tag_base, inst_id, step, beta, E, s2r,
                                                       att, succ, x.key, lane);
        wave_sync();
        return r;
    }
    const int nhalf = pilot ? 1 : 2;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const uint32_t c3s = ctr3(kTagShuf | tag_base, inst_id);
    const uint32_t c3e = ctr3(kTagExch | tag_base, inst_id);
    unsigned long long nsw = 0;
    for (int h = 0; h < nhalf; ++h) {
        // pair order of this half: indices sorted by (key, index)
        for (int j = lane; j < np; j += kWave) {
            x.key[j] = pick(philox4x32_10((uint32_t)j >> 2, (uint32_t)h, step, c3s, k0, k1), (uint32_t)j);
        }
        for (int k = lane; k < nrep; k += kWave) {
            x.tpos[k] = -1;
            x.level[k] = 0;
        }
        wave_sync();
        for (int j = lane; j < np; j += kWave) {
            const uint32_t kj = x.key[j];
            int rank = 0;
            for (int i = 0; i < np; ++i) {
                const uint32_t ki = x.key[i];
                rank += (ki < kj) || (ki == kj && i < j);
            }
            x.perm[rank] = j;
        }
        // log-uniform of every attempt of this half (attempt index a is global in the step)
        for (int i = lane; i < np; i += kWave) {
            const int a = h * np + i;
            x.logu[a] = uniform_log(pick(philox4x32_10((uint32_t)a >> 2, 0u, step, c3e, k0, k1), (uint32_t)a));
        }
        wave_sync();
        for (int i = lane; i < np; i += kWave) {
            const int a = h * np + i;
            const int chain = pilot ? 0 : (a & 1);
            x.tpos[x.perm[i] + chain * nt] = a;
        }
        wave_sync();
        // dependency levels: level(k) = 1 + max level of an earlier attempt on k-1 or k+1
        bool changed = true;
        while (changed) {
            bool mine = false;
            for (int k = lane; k < nrep; k += kWave) {
                const int tk = x.tpos[k];
                if (tk < 0) continue;
                const int kc = k % nt;  // index inside the chain; pairs are 0..nt-2
                int lv = 1;
                if (kc > 0) {
                    const int tl = x.tpos[k - 1];
                    if (tl >= 0 && tl < tk) lv = max(lv, x.level[k - 1] + 1);
                }
                if (kc + 1 < np) {
                    const int tr = x.tpos[k + 1];
                    if (tr >= 0 && tr < tk) lv = max(lv, x.level[k + 1] + 1);
                }
                if (lv != x.level[k]) {
                    x.level[k] = lv;
                    mine = true;
                }
            }
            wave_sync();
            changed = __ballot(mine) != 0ull;
        }
        int maxlv = 0;
        for (int k = lane; k < nrep; k += kWave) maxlv = max(maxlv, x.level[k]);
        maxlv = wave_max_i32(maxlv);
        for (int L = 1; L <= maxlv; ++L) {
            for (int k = lane; k < nrep; k += kWave) {
                if (x.level[k] != L) continue;
                const int r1 = s2r[k], r2 = s2r[k + 1];
                const double lb = (beta[k] - beta[k + 1]) * (double)(E[r1] - E[r2]);
                att[k] += 1u;
                if (x.logu[x.tpos[k]] < lb) {
                    s2r[k] = r2;
                    s2r[k + 1] = r1;
                    succ[k] += 1u;
                    ++nsw;
                }
            }
            wave_sync();
        }
    }
    return nsw;
}

// ---------------------------------------------------------------------------
// Houdayer cluster move on one slot pair (icm.cc:8-215), by one wavefront
// (Block = false) or by a whole workgroup (Block = true; tid/nthr span the
// workgroup).  S1/S2 hold both states (bit words, copied in from global), D
// the disagreement set, V the cluster; Q is a queue of up to n/2 positions.
// The cluster is grown breadth-first from a uniform seed in D; the reference
// grows it depth-first (icm.cc:37-72), which gives the same component.  With
// several waves the queue order depends on timing, but the component, the
// flipped states and the (integer) energy changes do not.  Returns the
// energy changes of both replicas; the new states stay in S1/S2 and are
// written to g1/g2 unless write_back is false.
// ---------------------------------------------------------------------------
struct IcmScratch {
    uint32_t* S1;
    uint32_t* S2;
    uint32_t* D;
    uint32_t* V;
    int32_t* tail;
    uint16_t* Q;
    long long* red;  // Block: one slot per wave for workgroup sums
    uint32_t* Do;    // lattice growth: D and the cluster in lattice order (LatView::words each)
    uint32_t* Vo;
};

// A periodic L^dim lattice (dim 2 or 3) whose sites i = x + L*y [+ L*L*z]
// are numbered like its rows: bit sets in lattice order keep each row of L
// sites in R = ceil(L/32) words, and a cluster grows by whole words --
// shifts along x, word moves along y and z.  wpos[32*w + b] is the position
// of bit b of word w (padding bits name position 0 and are masked), obit[p]
// the word << 5 | bit of position p.  dim = 0: no lattice (queue growth).
struct LatView {
    const uint16_t* wpos;
    const uint32_t* obit;
    int dim, L, R, words;
    const uint32_t* jn = nullptr;  // edge signs [dim][rows] (rows of one word), or none
    int half = 0;  // n / 2 when the colour classes are the site parities in index order (InstDesc::lsw_dim):
                   // a lattice word is then the bit interleave of two 16-bit class fields
};

// 16 bits to the even bits of a word, and back
__device__ __forceinline__ uint32_t spread16(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    return (x | (x << 1)) & 0x55555555u;
}
__device__ __forceinline__ uint32_t compress16(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    return (x | (x >> 8)) & 0x0000FFFFu;
}
// 32 bits of a position-order bit set from bit p (0 <= p < n; the set has a
// word past its last)
__device__ __forceinline__ uint32_t bits_at(const uint32_t* S, int p) {
    const int w0 = p >> 5;
    return (uint32_t)((((uint64_t)S[w0 + 1] << 32) | S[w0]) >> (p & 31));
}
__device__ __forceinline__ void bits_or(uint32_t* S, int p, uint32_t v) {
    if (!v) return;
    const uint64_t x = (uint64_t)v << (p & 31);
    const int w0 = p >> 5;
    if ((uint32_t)x) atomicOr(S + w0, (uint32_t)x);
    if ((uint32_t)(x >> 32)) atomicOr(S + w0 + 1, (uint32_t)(x >> 32));
}

// lattice words per lane kept in registers during cluster growth
constexpr int kLatSlots = 4;

__device__ __forceinline__ LatView lat_view(const DevArgs& a, const InstDesc& d) {
    if (d.lat_dim == 0) return LatView{nullptr, nullptr, 0, 0, 0, 0};
    const uint32_t* t = a.lat + d.lat_off;  // wpos (16-bit) then obit
    return LatView{reinterpret_cast<const uint16_t*>(t), t + 16 * (int64_t)d.lat_words, d.lat_dim, d.lat_L, d.lat_R,
                   d.lat_words, d.jn_off >= 0 ? a.jneg + d.jn_off : nullptr, d.lsw_dim > 0 ? d.n / 2 : 0};
}

// Grows the cluster of seed_pos (a position in D) as a bit set in lattice
// order until no word changes, then writes it to V (position order) and its
// positions to Q (with edge signs instead S1 in lattice order to the Q area,
// for lattice_cut_energy); returns the cluster size.  One wavefront.  Updates are
// in place and only ever add sites reached from the seed through D, so the
// order of the word updates does not change the fixed point: the connected
// component of the seed in D.
__device__ __forceinline__ int lattice_grow(const LatView& lv, int n, int nw, int seed_pos, IcmScratch sc,
                                            int lane) {
    const int L = lv.L, R = lv.R, NW = lv.words;
    const int lb = L - 32 * (R - 1);  // sites in a row's last word (1..32)
    const uint32_t lastmask = lb == 32 ? 0xFFFFFFFFu : ((1u << lb) - 1u);
    uint32_t* Do = sc.Do;
    uint32_t* Vo = sc.Vo;
    const uint32_t* D = sc.D;
    // With edge signs (the cut energy on bit sets follows) S1 goes to lattice
    // order in the same pass, into the position-list area, and no position
    // list is written.
    const bool cut = lv.jn != nullptr;
    uint32_t* S1o = reinterpret_cast<uint32_t*>(sc.Q);
    // the row's class-c sites x = 2i + par_c, positions c * half + row * H + i
    auto row_parity = [&](int row) {
        const int y = row % L;
        return lv.dim == 3 ? (y + row / L) & 1 : y & 1;
    };
    if (lv.half > 0) {
        // D (and S1) in lattice order from the classes' 16-bit fields: word k
        // of a row holds sites i in [16 k, 16 k + 16) of both classes,
        // interleaved by the row's parity
        const int H = L >> 1;
        for (int w = lane; w < NW; w += kWave) {
            const int row = w / R, k = w - row * R;
            const int par0 = row_parity(row);
            const int i0 = 16 * k, q = row * H + i0;
            const uint32_t fm = (1u << min(16, H - i0)) - 1u;
            const uint32_t e0 = spread16(bits_at(D, q) & fm), e1 = spread16(bits_at(D, lv.half + q) & fm);
            Do[w] = par0 ? (e1 | (e0 << 1)) : (e0 | (e1 << 1));
            Vo[w] = 0u;
            if (cut) {
                const uint32_t t0 = spread16(bits_at(sc.S1, q) & fm), t1 = spread16(bits_at(sc.S1, lv.half + q) & fm);
                S1o[w] = par0 ? (t1 | (t0 << 1)) : (t0 | (t1 << 1));
            }
        }
    } else
    // D in lattice order: the 32 positions of each word are read as four
    // 16-byte loads, then their D bits
    for (int w = lane; w < NW; w += kWave) {
        const uint4* src = reinterpret_cast<const uint4*>(lv.wpos + 32 * w);
        uint4 c[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) c[q] = src[q];
        uint32_t word = 0, sword = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t pv[4] = {c[q].x, c[q].y, c[q].z, c[q].w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                const uint32_t p0 = pv[h] & 0xFFFFu, p1 = pv[h] >> 16;
                word |= ((D[p0 >> 5] >> (p0 & 31)) & 1u) << (8 * q + 2 * h);
                word |= ((D[p1 >> 5] >> (p1 & 31)) & 1u) << (8 * q + 2 * h + 1);
                if (cut) {
                    sword |= ((sc.S1[p0 >> 5] >> (p0 & 31)) & 1u) << (8 * q + 2 * h);
                    sword |= ((sc.S1[p1 >> 5] >> (p1 & 31)) & 1u) << (8 * q + 2 * h + 1);
                }
            }
        }
        if (w % R == R - 1) {
            word &= lastmask;
            sword &= lastmask;
        }
        Do[w] = word;
        Vo[w] = 0u;
        if (cut) S1o[w] = sword;
    }
    wave_sync();
    if (lane == 0) {
        const uint32_t ob = lv.obit[seed_pos];
        Vo[ob >> 5] |= 1u << (ob & 31);
    }
    wave_sync();
    const int rows = NW / R;
    // neighbour rows of row r along y and z (wrapped within the lattice)
    auto rows_of = [&](int row, int& yu, int& yd, int& zu, int& zd) {
        const int y = row % L, zb = row - y;
        yu = zb + (y + 1 == L ? 0 : y + 1);
        yd = zb + (y == 0 ? L - 1 : y - 1);
        zu = zd = row;
        if (lv.dim == 3) {
            const int z = row / L;
            zu = y + L * (z + 1 == L ? 0 : z + 1);
            zd = y + L * (z == 0 ? L - 1 : z - 1);
        }
    };
    // x closure of a one-word row: bits reached along x through d (wrapped)
    auto close_x = [&](uint32_t nv, uint32_t d) {
        while (true) {
            const uint32_t t = nv | (((nv << 1) | (nv >> (L - 1)) | (nv >> 1) | (nv << (L - 1))) & d);
            if (t == nv) return nv;
            nv = t;
        }
    };
    if (R == 1 && rows <= kWave) {
        // one row per lane, all in registers: y and z neighbours by lane shuffles
        const bool own = lane < rows;
        const uint32_t d = own ? Do[lane] : 0u;
        uint32_t v = own ? Vo[lane] : 0u;
        int yu = lane, yd = lane, zu = lane, zd = lane;
        if (own) rows_of(lane, yu, yd, zu, zd);
        while (true) {
            uint32_t ny = __shfl(v, yu, kWave) | __shfl(v, yd, kWave);
            if (lv.dim == 3) ny |= __shfl(v, zu, kWave) | __shfl(v, zd, kWave);
            const uint32_t nv = close_x(v | (ny & d), d);
            const bool grew = nv != v;
            v = nv;
            if (__ballot(grew) == 0ull) break;
        }
        if (own) Vo[lane] = v;
        wave_sync();
    } else if (lv.dim == 2 && R <= 4 && rows <= 2 * kWave) {
        // 2D, rows of up to 4 words and up to 128 rows: a lane holds rows
        // lane and lane + 64 whole, in registers.  A row closes along x in
        // registers (fill_row: additions carried across its words, wrapped
        // at bit L-1); rows y +- 1 come by DPP wave shifts and readlane.
        uint32_t d[2][4], v[2][4];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int row = lane + kWave * j;
            const bool own = row < rows;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const bool in = own && k < R;
                d[j][k] = in ? Do[row * R + k] : 0u;
                v[j][k] = in ? Vo[row * R + k] : 0u;
            }
        }
        const bool two = rows > kWave;  // rows 64.. in slot 1
        const int last = rows - 1;
        // the row's x closure within d (wrapped), by fill_row
        auto close_row = [&](uint32_t* rv, const uint32_t* rd) {
            switch (R) {
                case 1: fill_row<1>(rv, rd, L); break;
                case 2: fill_row<2>(rv, rd, L); break;
                case 3: fill_row<3>(rv, rd, L); break;
                default: fill_row<4>(rv, rd, L); break;
            }
        };
        while (true) {
            bool grew = false;
            uint32_t nv[2][4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k >= R) {  // wave-uniform
                    nv[0][k] = nv[1][k] = 0u;
                    continue;
                }
                // the same word of rows y +- 1: lane +- 1 of the same slot by
                // DPP wave shifts; the rows across a slot or the wrap
                // (0, 63, 64, rows - 1) by readlane
                const uint32_t a0 = v[0][k], a1 = v[1][k];
                const uint32_t r0 = __builtin_amdgcn_readlane(a0, 0);
                const uint32_t rl = __builtin_amdgcn_readlane(last >= kWave ? a1 : a0, last & (kWave - 1));
                const uint32_t n0 = (uint32_t)lane_next((int)a0), p0 = (uint32_t)lane_prev((int)a0);
                uint32_t u = lane == last ? r0 : n0;
                if (two && lane == kWave - 1) u = __builtin_amdgcn_readlane(a1, 0);
                const uint32_t w = lane == 0 ? rl : p0;
                nv[0][k] = a0 | ((u | w) & d[0][k]);
                if (two) {
                    const uint32_t r63 = __builtin_amdgcn_readlane(a0, kWave - 1);
                    const uint32_t n1 = (uint32_t)lane_next((int)a1), p1 = (uint32_t)lane_prev((int)a1);
                    const uint32_t u1 = lane + kWave == last ? r0 : n1;
                    const uint32_t w1 = lane == 0 ? r63 : p1;
                    nv[1][k] = a1 | ((u1 | w1) & d[1][k]);
                } else {
                    nv[1][k] = 0u;
                }
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (j == 1 && !two) break;  // slot 1 stays empty
                close_row(nv[j], d[j]);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    grew |= nv[j][k] != v[j][k];
                    v[j][k] = nv[j][k];
                }
            }
            if (__ballot(grew) == 0ull) break;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int row = lane + kWave * j;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (row < rows && k < R) Vo[row * R + k] = v[j][k];
        }
        wave_sync();
    } else if (NW <= kLatSlots * kWave) {
        // up to kLatSlots words per lane kept in registers, updated in place
        // slot by slot (monotone: a neighbour read before or after its own
        // update this round only changes when a site joins, not whether), so
        // a later slot already sees the earlier slots' growth; one wave
        // barrier per round
        const int J = (NW + kWave - 1) / kWave;
        uint32_t dreg[kLatSlots], vreg[kLatSlots];
#pragma unroll
        for (int j = 0; j < kLatSlots; ++j) {
            const int w = lane + kWave * j;
            const bool own = j < J && w < NW;
            dreg[j] = own ? Do[w] : 0u;
            vreg[j] = own ? Vo[w] : 0u;
        }
        while (true) {
            bool grew = false;
#pragma unroll
            for (int j = 0; j < kLatSlots; ++j) {
                const int w = lane + kWave * j;
                if (j >= J || w >= NW || vreg[j] == dreg[j]) continue;
                const int row = w / R, k = w - row * R;
                int yu, yd, zu, zd;
                rows_of(row, yu, yd, zu, zd);
                uint32_t n = Vo[yu * R + k] | Vo[yd * R + k];
                if (lv.dim == 3) n |= Vo[zu * R + k] | Vo[zd * R + k];
                const uint32_t v = vreg[j];
                if (R == 1) {
                    n |= ((v << 1) | (v >> (L - 1)) | (v >> 1) | (v << (L - 1))) & lastmask;
                } else {
                    const int rw = row * R;
                    const uint32_t cin = k > 0 ? (Vo[w - 1] >> 31) : ((Vo[rw + R - 1] >> (lb - 1)) & 1u);
                    const uint32_t cout = k < R - 1 ? ((Vo[w + 1] & 1u) << 31) : ((Vo[rw] & 1u) << (lb - 1));
                    uint32_t nx = (v << 1) | cin | (v >> 1) | cout;
                    if (k == R - 1) nx &= lastmask;
                    n |= nx;
                }
                uint32_t nv = v | (n & dreg[j]);
                if (R == 1) nv = close_x(nv, dreg[j]);
                if (nv != v) {
                    vreg[j] = nv;
                    Vo[w] = nv;
                    grew = true;
                }
            }
            wave_sync();
            if (__ballot(grew) == 0ull) break;
        }
    } else {
        // larger lattices: words updated in place (monotone, any order)
        while (true) {
            bool grew = false;
            for (int w = lane; w < NW; w += kWave) {
                const uint32_t d = Do[w];
                const uint32_t v = Vo[w];
                if (v == d) continue;  // every site of D in this word is in already
                const int row = w / R, k = w - row * R;
                uint32_t nx;
                if (R == 1) {
                    nx = ((v << 1) | (v >> (L - 1)) | (v >> 1) | (v << (L - 1))) & lastmask;
                } else {
                    const int rw = row * R;
                    const uint32_t cin = k > 0 ? (Vo[w - 1] >> 31) : ((Vo[rw + R - 1] >> (lb - 1)) & 1u);
                    const uint32_t cout = k < R - 1 ? ((Vo[w + 1] & 1u) << 31) : ((Vo[rw] & 1u) << (lb - 1));
                    nx = (v << 1) | cin | (v >> 1) | cout;
                    if (k == R - 1) nx &= lastmask;
                }
                int yu, yd, zu, zd;
                rows_of(row, yu, yd, zu, zd);
                uint32_t ny = Vo[yu * R + k] | Vo[yd * R + k];
                if (lv.dim == 3) ny |= Vo[zu * R + k] | Vo[zd * R + k];
                uint32_t nv = v | ((nx | ny) & d);
                if (R == 1) nv = close_x(nv, d);
                if (nv != v) {
                    Vo[w] = nv;
                    grew = true;
                }
            }
            wave_sync();
            if (__ballot(grew) == 0ull) break;
        }
    }
    int csize = 0;
    if (lv.half > 0 && cut) {
        // back to position order by the classes' fields (no position list
        // follows the cut energy on bit sets)
        for (int w = lane; w < nw; w += kWave) sc.V[w] = 0u;
        wave_sync();
        const int H = L >> 1;
        int cs = 0;
        for (int w = lane; w < NW; w += kWave) {
            const uint32_t v = Vo[w];
            cs += __popc(v);
            if (v) {
                const int row = w / R, k = w - row * R;
                const int par0 = row_parity(row);
                const uint32_t fe = compress16(v), fo = compress16(v >> 1);
                const int q = row * H + 16 * k;
                bits_or(sc.V, q, par0 ? fo : fe);
                bits_or(sc.V, lv.half + q, par0 ? fe : fo);
            }
        }
        csize = wave_sum_i32(cs);
        wave_sync();
        return csize;
    }
    // back to position order: V, and the cluster's positions in Q (the
    // table reads of four 64-position blocks are issued together)
    for (int p00 = 0; p00 < 32 * nw; p00 += 4 * kWave) {
        uint32_t ob[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = p00 + u * kWave + lane;
            ob[u] = p < n ? lv.obit[p] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p0 = p00 + u * kWave;
            if (p0 >= 32 * nw) break;
            const int p = p0 + lane;
            const bool in = p < n && ((Vo[ob[u] >> 5] >> (ob[u] & 31)) & 1u) != 0u;
            const unsigned long long b = __ballot(in);
            const int below =
                (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
            if (in && !cut) sc.Q[csize + below] = (uint16_t)p;
            if (lane == 0) sc.V[p0 >> 5] = (uint32_t)b;
            if (lane == 32 && (p0 >> 5) + 1 < nw) sc.V[(p0 >> 5) + 1] = (uint32_t)(b >> 32);
            csize += __popcll(b);
        }
    }
    wave_sync();
    return csize;
}

// Energy change of flipping the grown cluster (Vo, lattice order) in both
// replicas of a periodic lattice with no field, on bit sets (one wavefront,
// after lattice_grow).  An edge (p, q) with exactly one end in the cluster
// is cut; flipping the cluster turns its w s_p s_q = (-1)^(neg_pq ^ b_p ^
// b_q) into the negative, so dE = -2 (cut - 2 odd), odd the cut edges whose
// parity bit is 1 (replica 1; replica 2's change is its negative, as the
// cluster is a component of D).  Each edge is counted once, from the site to its +axis
// neighbour: along x the row rotated by one site (across its R words, the
// last word's top site taking the first word's bit 0), along y and z the
// same word of the next row, wrapped.  A lane takes whole rows.
// lattice_grow left S1 in lattice order in the position-list area.  Adds
// this lane's part to dE1 and its negative to dE2.
__device__ __forceinline__ void lattice_cut_energy(const LatView& lv, IcmScratch sc, int lane, long long& dE1,
                                                   long long& dE2) {
    const int L = lv.L, R = lv.R, NW = lv.words, rows = NW / R;
    const int lb = L - 32 * (R - 1);  // sites in a row's last word (1..32)
    const uint32_t lastmask = lb == 32 ? 0xFFFFFFFFu : ((1u << lb) - 1u);
    const uint32_t* Vo = sc.Vo;
    const uint32_t* S1o = reinterpret_cast<const uint32_t*>(sc.Q);
    const uint32_t* jx = lv.jn;
    const uint32_t* jy = jx + NW;
    const uint32_t* jz = jy + NW;
    int cut = 0, odd = 0;
    for (int row = lane; row < rows; row += kWave) {
        const int y = row % L, zb = row - y;
        const int ry = zb + (y + 1) % L;
        const int rz = lv.dim == 3 ? y + (zb + L) % (L * L) : 0;
        const int w0 = row * R;
        for (int k = 0; k < R; ++k) {
            const int w = w0 + k;
            const uint32_t c = Vo[w], s1 = S1o[w];
            auto edges = [&](uint32_t cn, uint32_t s1n, uint32_t neg) {
                const uint32_t x = c ^ cn;
                cut += __popc(x);
                odd += __popc(x & (neg ^ s1 ^ s1n));
            };
            // +x: bit b of the rotated word is site 32 k + b + 1 of the row
            const bool last = k == R - 1;
            const int wn = last ? w0 : w + 1;
            const uint32_t cn = Vo[wn], s1n = S1o[wn];
            const int top = last ? lb - 1 : 31;
            const uint32_t m = last ? lastmask : 0xFFFFFFFFu;
            edges(((c >> 1) | ((cn & 1u) << top)) & m, ((s1 >> 1) | ((s1n & 1u) << top)) & m, jx[w]);
            const int wy = ry * R + k;
            edges(Vo[wy], S1o[wy], jy[w]);
            if (lv.dim == 3) {
                const int wz = rz * R + k;
                edges(Vo[wz], S1o[wz], jz[w]);
            }
        }
    }
    const long long t = -2ll * (cut - 2 * odd);
    dE1 += t;
    dE2 -= t;
}

// Per-wave slots of the workgroup reductions in the ICM team code.
constexpr int kIcmRedSlots = 16;
// Clusters of graphs with at least this many neighbours per spin on average
// (and no lattice form) grow with a whole workgroup where one is at hand:
// their breadth-first levels are wide (an ER graph of mean degree 20 reaches
// a few hundred sites in 4-5 levels), and one wave would issue every
// instruction of them alone.
constexpr int kIcmBlockDeg = 8;

template <bool Block>
__device__ __forceinline__ void team_sync() {
    if (Block)
        __syncthreads();
    else
        wave_sync();
}

template <bool Block>
__device__ __forceinline__ long long team_sum_i64(long long v, long long* red, int tid, int nthr) {
    v = wave_sum_i64(v);
    if (!Block || nthr <= kWave) return v;
    __syncthreads();  // the previous sum's reads of red are done
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = v;
    __syncthreads();
    long long s = 0;
    for (int q = 0; q < nthr / kWave; ++q) s += red[q];
    return s;
}

template <bool Block>
__device__ __forceinline__ int team_sum_i32(int v, long long* red, int tid, int nthr) {
    v = wave_sum_i32(v);
    if (!Block || nthr <= kWave) return v;
    __syncthreads();
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = v;
    __syncthreads();
    long long s = 0;
    for (int q = 0; q < nthr / kWave; ++q) s += red[q];
    return (int)s;
}

template <int Fmt, bool Block>
__device__ __forceinline__ void icm_team(const Csr<Fmt>& csr, const int32_t* hf, bool has_field, int n,
                                         int nw, int t, uint32_t step, uint64_t seed, uint32_t c3,
                                         const uint32_t* g1_in, const uint32_t* g2_in, IcmScratch sc,
                                         long long& dE1_out, long long& dE2_out, int tid, int nthr,
                                         bool write_back = true, LatView lv = LatView{nullptr, nullptr, 0, 0, 0, 0},
                                         bool have_e = false, long long e1 = 0, long long e2 = 0) {
    // the new states are left in sc.S1/sc.S2 and, with write_back, stored to g1/g2;
    // have_e: e1, e2 are the two replicas' energies (queue growth may then
    // finish bottom-up and take the energy change from the rest of D)
    uint32_t* g1 = const_cast<uint32_t*>(g1_in);
    uint32_t* g2 = const_cast<uint32_t*>(g2_in);
    uint32_t* S1 = sc.S1;
    uint32_t* S2 = sc.S2;
    uint32_t* D = sc.D;
    uint32_t* V = sc.V;
    const int lane = tid & (kWave - 1);
    const uint32_t lastmask = (n & 31) ? ((1u << (n & 31)) - 1u) : 0xFFFFFFFFu;
    int cnt = 0;
    for (int w = tid; w < nw; w += nthr) {
        const uint32_t m = (w == nw - 1) ? lastmask : 0xFFFFFFFFu;
        const uint32_t x1 = g1[w] & m, x2 = g2[w] & m;
        S1[w] = x1;
        S2[w] = x2;
        D[w] = x1 ^ x2;
        V[w] = 0u;
        cnt += __popc(x1 ^ x2);
    }
    cnt = team_sum_i32<Block>(cnt, sc.red, tid, nthr);
    long long dE1 = 0, dE2 = 0;
    team_sync<Block>();
    if (cnt > n / 2) {  // icm.cc:20-28: flip replica 1 globally
        if (has_field) {
            for (int p = tid; p < n; p += nthr) {
                const int s1 = (S1[p >> 5] >> (p & 31)) & 1u;
                dE1 += -2ll * (s1 ? 1 : -1) * hf[p];
            }
        }
        team_sync<Block>();
        for (int w = tid; w < nw; w += nthr) {
            const uint32_t m = (w == nw - 1) ? lastmask : 0xFFFFFFFFu;
            S1[w] = ~S1[w] & m;
            D[w] = ~D[w] & m;
        }
        cnt = n - cnt;
        team_sync<Block>();
    }
    if (cnt > 0) {
        int seed_pos = -1;
        if (tid < kWave) {  // the first wave picks the seed
            const Philox4 x = philox4x32_10((uint32_t)t >> 2, 0u, step, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
            const int k = (int)(((unsigned long long)pick(x, (uint32_t)t) * (unsigned long long)cnt) >> 32);
            // k-th element of D in position order: per-64-word chunks, inclusive scan
            int base = 0;
            for (int w0 = 0; w0 < nw && seed_pos < 0; w0 += kWave) {
                const int w = w0 + lane;
                const uint32_t dw = w < nw ? D[w] : 0u;
                const int c = __popc(dw);
                const int incl = wave_prefix_i32(c);  // DPP scan, no LDS crossbar
                const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
                if (k < base + total) {
                    const bool mine = (k < base + incl) && (k >= base + incl - c);
                    const unsigned long long ball = __ballot(mine);
                    const int owner = __ffsll((long long)ball) - 1;
                    int sel = -1;
                    if (lane == owner) {
                        // the kk-th set bit of dw by halving: 5 popcounts, no loop over bits
                        int kk = k - (base + incl - c), at = 0;
                        uint32_t x = dw;
#pragma unroll
                        for (int sh = 16; sh >= 1; sh >>= 1) {
                            const uint32_t lo = x & ((1u << sh) - 1u);
                            const int cl = __popc(lo);
                            if (kk >= cl) {
                                kk -= cl;
                                x >>= sh;
                                at += sh;
                            } else {
                                x = lo;
                            }
                        }
                        sel = 32 * w + at;
                    }
                    seed_pos = __builtin_amdgcn_readlane(sel, owner);  // owner is wave-uniform
                }
                base += total;
            }
            if (lane == 0 && (Block || lv.dim == 0)) {
                V[seed_pos >> 5] |= 1u << (seed_pos & 31);
                sc.Q[0] = (uint16_t)seed_pos;
                if (Block) sc.tail[0] = 1;
            }
        }
        team_sync<Block>();
        constexpr int K = Csr<Fmt>::kPerChunk;
        int csize;
        bool cut_done = false;  // the energy change is known (bit sets, or sums taken during the growth)
        if (!Block && lv.dim > 0) {
            csize = lattice_grow(lv, n, nw, seed_pos, sc, lane);
            if (lv.jn) {
                lattice_cut_energy(lv, sc, lane, dE1, dE2);
                cut_done = true;
            }
        } else {
            // Breadth-first growth.  Each lane takes one queued site and walks its
            // row a 16-byte chunk at a time: the D and V words of the chunk's
            // neighbours are read together, candidates are claimed with atomicOr
            // on V, and the newly claimed sites are appended at wave-scanned
            // offsets.  One wave keeps the queue tail in a register; a workgroup
            // keeps it in LDS (one atomic add per wave and chunk) and reads it
            // back between two barriers at the end of each round.
            int head = 0, tail = 1;
            bool bottom_up = false;
            const long long gflip = dE1;  // this thread's field terms of a global flip of replica 1
            while (head < tail) {
                // once the queued sites are at least twice the D sites not yet
                // claimed, the rest of the component is found bottom-up (on
                // 512 ER G(5000, 0.004): 0.73 ms of cluster moves per step,
                // against 0.79 ms switching at half and 0.90 ms at an eighth)
                if (have_e && (tail - head) >= 2 * (cnt - tail)) {  // uniform
                    bottom_up = true;
                    break;
                }
                const int tl = tail;
                const int idx = head + tid;
                const bool act = idx < tl;
                const int p = act ? (int)sc.Q[idx] : 0;
                const int dg = act ? csr.deg(p) : 0;
                const uint8_t* row = csr.row(p);
                const int nch = (dg + K - 1) / K;
                // p's boundary sum, taken while its row is read: a neighbour
                // of a cluster site is outside the cluster exactly when it
                // is outside D (the cluster is D's component)
                int b1 = 0;
                // chunks up to the longest row among the wave's queued sites
                // (lanes without a site have none), not the instance's
                const int wch = wave_max_i32(nch);
                for (int ch = 0; ch < wch; ++ch) {
                    const int live = ch < nch ? min(K, dg - ch * K) : 0;
                    // A lane loads a chunk only while its row has one; the loads
                    // of a chunk's records are issued together so that they
                    // overlap, and liveness masks the padding records' results.
                    int col[K], wgt[K];
                    uint32_t dwd[K], vwd[K], swd[K];
                    if (live > 0) {
                        csr.chunk(row, ch, col, wgt);
    #pragma unroll
                        for (int h = 0; h < K; ++h) {
                            dwd[h] = D[col[h] >> 5];
                            vwd[h] = V[col[h] >> 5];
                            swd[h] = S1[col[h] >> 5];
                        }
                    } else {
    #pragma unroll
                        for (int h = 0; h < K; ++h) {
                            col[h] = 0;
                            wgt[h] = 0;
                            dwd[h] = vwd[h] = swd[h] = 0u;
                        }
                    }
                    uint32_t cand = 0;
    #pragma unroll
                    for (int h = 0; h < K; ++h) {
                        const uint32_t bit = 1u << (col[h] & 31);
                        const bool c = (h < live) & ((dwd[h] & bit) != 0u) & ((vwd[h] & bit) == 0u);
                        cand |= (c ? 1u : 0u) << h;
                        const bool out = (h < live) & ((dwd[h] & bit) == 0u);
                        b1 += out ? ((swd[h] & bit) ? wgt[h] : -wgt[h]) : 0;
                    }
                    // only candidates claim (padding records share one sentinel word,
                    // and same-address atomics serialise); the claims carry no
                    // dependence on each other, so they still issue back to back
                    uint32_t old[K];
    #pragma unroll
                    for (int h = 0; h < K; ++h) {
                        old[h] = 0xFFFFFFFFu;
                        if ((cand >> h) & 1u) old[h] = atomicOr(&V[col[h] >> 5], 1u << (col[h] & 31));
                    }
                    uint32_t claimed = 0;
    #pragma unroll
                    for (int h = 0; h < K; ++h)
                        if (((cand >> h) & 1u) && !(old[h] & (1u << (col[h] & 31)))) claimed |= 1u << h;
                    // claims of the lanes below this one (and of the whole wave)
                    // from one ballot per record slot: no cross-lane data moves
                    int excl = 0, total = 0;
    #pragma unroll
                    for (int h = 0; h < K; ++h) {
                        const unsigned long long b = __ballot((claimed >> h) & 1u);
                        excl += (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
                        total += __popcll(b);
                    }
                    int pos;
                    if (Block) {
                        int wbase = 0;
                        if (lane == 0 && total > 0) wbase = atomicAdd(sc.tail, total);
                        pos = __builtin_amdgcn_readfirstlane(wbase) + excl;
                    } else {
                        pos = tail + excl;
                        tail += total;
                    }
    #pragma unroll
                    for (int h = 0; h < K; ++h)
                        if ((claimed >> h) & 1u) sc.Q[pos++] = (uint16_t)col[h];
                }
                if (act) {
                    const int s1 = ((S1[p >> 5] >> (p & 31)) & 1u) ? 1 : -1;
                    const long long t = -2ll * s1 * (b1 + (has_field ? hf[p] : 0));
                    dE1 += t;
                    dE2 -= t;
                }
                head = (tl - head > nthr) ? head + nthr : tl;
                team_sync<Block>();
                if (Block) {
                    tail = sc.tail[0];
                    __syncthreads();  // every wave has the tail before the next round appends
                }
            }
            csize = tail;
            cut_done = true;  // boundary sums taken during the growth
            if (bottom_up) {
                // Bottom-up passes: every unclaimed site of D with a claimed
                // neighbour is claimed (a row is read only up to its first
                // claimed neighbour), until a pass claims none.  Claimed sites
                // are in the seed's component of D and the component is closed
                // under D-adjacency, so V ends as exactly that component, the
                // same set the breadth-first growth finds.
                int added;
                do {
                    int mine = 0;
                    for (int p = tid; p < n; p += nthr) {
                        const uint32_t bit = 1u << (p & 31);
                        if ((D[p >> 5] & bit) == 0u || (V[p >> 5] & bit) != 0u) continue;
                        const uint8_t* row = csr.row(p);
                        const int dg = csr.deg(p);
                        bool hit = false;
                        for (int ch = 0; K * ch < dg && !hit; ++ch) {
                            int col[K], wgt[K];
                            csr.chunk(row, ch, col, wgt);
                            const int live = min(K, dg - ch * K);
#pragma unroll
                            for (int h = 0; h < K; ++h)
                                hit = hit || (h < live && ((V[col[h] >> 5] >> (col[h] & 31)) & 1u) != 0u);
                        }
                        if (hit) {
                            atomicOr(&V[p >> 5], bit);
                            ++mine;
                        }
                    }
                    team_sync<Block>();  // this pass's claims are visible to the next
                    added = team_sum_i32<Block>(mine, sc.red, tid, nthr);
                } while (added > 0);
                // Energy change from the rest R = D \ C.  Flipping C in
                // replica 1 (after its global flip, if any) gives replica 2's
                // state with R flipped (on C s1 becomes s2, on R s1 = -s2,
                // elsewhere s1 = s2), so dE1 = e2 - e1 + (change of flipping R
                // in replica 2), global flip included.  Replica 2 ends as
                // replica 1 after its global flip with R flipped, and R is a
                // union of components of D, so dE2 = -dE1 + (the global
                // flip's change).  The boundary sums taken for the sites
                // grown top-down are dropped.
                dE1 = tid == 0 ? e2 - e1 : 0;
                for (int p = tid; p < n; p += nthr) {
                    const uint32_t bit = 1u << (p & 31);
                    if ((D[p >> 5] & bit) == 0u || (V[p >> 5] & bit) != 0u) continue;  // p in R
                    const int s2 = ((S2[p >> 5] & bit) != 0u) ? 1 : -1;
                    const uint8_t* row = csr.row(p);
                    const int dg = csr.deg(p);
                    int b2 = 0;  // sum over neighbours j outside R of J s2_j
                    for (int ch = 0; K * ch < dg; ++ch) {
                        int col[K], wgt[K];
                        csr.chunk(row, ch, col, wgt);
                        const int live = min(K, dg - ch * K);
#pragma unroll
                        for (int h = 0; h < K; ++h) {
                            const int j = col[h];
                            const uint32_t sh = (uint32_t)(j & 31);
                            const bool in_r = ((D[j >> 5] >> sh) & 1u) != 0u && ((V[j >> 5] >> sh) & 1u) == 0u;
                            const int a2 = ((S2[j >> 5] >> sh) & 1u) ? wgt[h] : -wgt[h];
                            b2 += (h < live && !in_r) ? a2 : 0;
                        }
                    }
                    dE1 += -2ll * s2 * (b2 + (has_field ? hf[p] : 0));
                }
                dE2 = gflip - dE1;
            }
        }
        // energy change of flipping the cluster C in replica 1: boundary
        // edges (p in C, j not in C) and the field terms of C.  Replica 2's
        // is its negative: C is a component of D = S1 ^ S2, so s2 = -s1 on C
        // and every neighbour of C outside C lies outside D, where s2 = s1.
        for (int idx = tid; idx < (cut_done ? 0 : csize); idx += nthr) {
            const int p = sc.Q[idx];
            const int s1 = ((S1[p >> 5] >> (p & 31)) & 1u) ? 1 : -1;
            int b1 = 0;
            const uint8_t* row = csr.row(p);
            const int dg = csr.deg(p);
            for (int ch = 0; K * ch < dg; ++ch) {
                int col[K], wgt[K];
                csr.chunk(row, ch, col, wgt);
                const int live = min(K, dg - ch * K);
#pragma unroll
                for (int h = 0; h < K; ++h) {
                    const int j = col[h];
                    const uint32_t sh = (uint32_t)(j & 31);
                    const uint32_t vin = (V[j >> 5] >> sh) & 1u;
                    const int a1 = ((S1[j >> 5] >> sh) & 1u) ? wgt[h] : -wgt[h];
                    const bool use = h < live && !vin;
                    b1 += use ? a1 : 0;
                }
            }
            const long long t = -2ll * s1 * (b1 + (has_field ? hf[p] : 0));
            dE1 += t;
            dE2 -= t;
        }
        team_sync<Block>();
        for (int w = tid; w < nw; w += nthr) {
            S1[w] ^= V[w];
            S2[w] ^= V[w];
        }
        team_sync<Block>();
    }
    dE1_out = team_sum_i64<Block>(dE1, sc.red, tid, nthr);
    dE2_out = team_sum_i64<Block>(dE2, sc.red, tid, nthr);
    if (write_back) {
        for (int w = tid; w < nw; w += nthr) {
            g1[w] = S1[w];
            g2[w] = S2[w];
        }
    }
}

// LDS bytes of the ICM scratch for n spins (S1,S2,D,V words, tail, queue of n/2+2).
// The four word arrays have nw + 1 entries so that padding records (which
// name position 32 * nw) stay inside them.
__host__ __device__ inline int icm_scratch_bytes(int n) {
    const int nw = (n + 31) / 32;
    return 16 * (nw + 1) + 16 + 2 * (n / 2 + 2);
}

}  // namespace ptx
