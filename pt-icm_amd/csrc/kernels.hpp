// kernels.hpp -- device-side building blocks shared by the split kernels
// (one launch per phase) and the fused per-instance kernel.  One wavefront
// (64 lanes) owns one replica during a sweep; spins of that replica are
// expanded to one byte per spin in LDS so neighbour reads are single
// ds_read_u8 and flips are plain byte stores (no atomics: inside a colour
// class no two updated sites are adjacent).
#pragma once

#include <hip/hip_runtime.h>

#include "ptx_common.h"

namespace ptx {

constexpr int kWave = 64;

// Arguments common to every kernel (device pointers of the engine).
struct DevArgs {
    const InstDesc* desc;
    const int32_t* rowptr;   // absolute offsets into adj
    const uint32_t* adj;     // (neighbour position) | (weight << 16)
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
    const int2* items;       // (instance, first replica) per sweep workgroup
    const int2* pairs;       // (instance, slot t) per ICM wavefront
    uint64_t seed;
    uint32_t tag_base;       // 0 or kTagPilot
    int32_t sweeps;
    int32_t spin_stride;     // LDS bytes per wave for spin bytes
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

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Bytes <-> bits.  A state word covers 32 consecutive positions; the byte
// buffer has nw*32 entries (padding bytes stay 0).
__device__ __forceinline__ void expand_bits(const uint32_t* __restrict__ words, int nw, uint8_t* sp, int lane) {
    for (int w = lane; w < nw; w += kWave) {
        const uint32_t v = words[w];
        uint32_t* dst = reinterpret_cast<uint32_t*>(sp + 32 * w);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint32_t x = (v >> (4 * q)) & 0xFu;
            dst[q] = (x | (x << 7) | (x << 14) | (x << 21)) & 0x01010101u;
        }
    }
}
__device__ __forceinline__ void pack_bits(const uint8_t* sp, int nw, uint32_t* __restrict__ words, int lane) {
    for (int w = lane; w < nw; w += kWave) {
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

// One Metropolis sweep of one replica in colour-class order
// (src/replica.cc:5-38 restated for class-parallel updates).  Each lane takes
// aligned groups of 4 positions, draws one Philox block per group (word k for
// position 4g+k), computes the local fields from the byte spins in LDS and
// accepts dE <= 0 always, dE > 0 iff draw < thr[dE/2 - 1].  After every class
// that accepted a flip the replica's best energy is updated
// (replica.cc:23-27 at class granularity).
struct SweepState {
    int32_t E, bestE, best_sweep;
    unsigned long long flips;
    bool best_dirty;
};

__device__ __forceinline__ void sweep_replica(const DevArgs& a, const InstDesc& d, int r, uint32_t gsweep, uint8_t* sp,
                                              const uint32_t* thr_row, uint32_t* best_words, SweepState& st,
                                              int lane) {
    const uint32_t c3 = ctr3(kTagSweep | a.tag_base, d.inst_id);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const int32_t* __restrict__ cls = a.cls + d.cls_off;
    const int32_t* __restrict__ rp = a.rowptr + d.rowptr_off;
    const int32_t* __restrict__ hf = a.hfield + d.h_off;
    const uint32_t* __restrict__ adj = a.adj;
    for (int c = 0; c < d.ncls; ++c) {
        const int cs = cls[c], ce = cls[c + 1];
        int dsum = 0, nacc = 0;
        for (int g = (cs >> 2) + lane; 4 * g < ce; g += kWave) {
            const int base = 4 * g;
            int dE[4];
            bool any_up = false;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int p = base + k;
                dE[k] = 0;
                if (p >= cs && p < ce) {
                    int f = d.has_field ? hf[p] : 0;
                    const int q0 = rp[p], q1 = rp[p + 1];
                    for (int q = q0; q < q1; ++q) {
                        const uint32_t rec = adj[q];
                        const int w = (int32_t)rec >> 16;
                        f += sp[rec & 0xFFFFu] ? w : -w;
                    }
                    dE[k] = sp[p] ? -2 * f : 2 * f;
                    any_up |= dE[k] > 0;
                }
            }
            Philox4 rnd = {{0u, 0u, 0u, 0u}};
            if (any_up) rnd = philox4x32_10((uint32_t)g, (uint32_t)r, gsweep, c3, k0, k1);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int p = base + k;
                if (p >= cs && p < ce) {
                    const bool acc = dE[k] <= 0 || rnd.v[k] < thr_row[(dE[k] >> 1) - 1];
                    if (acc) {
                        sp[p] ^= 1u;
                        dsum += dE[k];
                        ++nacc;
                    }
                }
            }
        }
        wave_sync();
        dsum = wave_sum_i32(dsum);
        nacc = wave_sum_i32(nacc);
        st.E += dsum;
        st.flips += (unsigned long long)nacc;
        if (nacc > 0 && st.E < st.bestE) {
            st.bestE = st.E;
            st.best_sweep = (int32_t)gsweep;
            pack_bits(sp, d.nw, best_words, lane);
        }
    }
}

// Energy of the byte state in LDS: E = sum_p h_p s_p + 1/2 sum_p s_p sum_q w s_q
// (src/problem.cc:121-142).
__device__ __forceinline__ int32_t energy_bytes(const DevArgs& a, const InstDesc& d, const uint8_t* sp, int lane) {
    const int32_t* __restrict__ rp = a.rowptr + d.rowptr_off;
    const int32_t* __restrict__ hf = a.hfield + d.h_off;
    long long acc2 = 0;
    for (int p = lane; p < d.n; p += kWave) {
        const int sgn = sp[p] ? 1 : -1;
        int f = 0;
        for (int q = rp[p]; q < rp[p + 1]; ++q) {
            const uint32_t rec = a.adj[q];
            const int w = (int32_t)rec >> 16;
            f += sp[rec & 0xFFFFu] ? w : -w;
        }
        acc2 += (long long)sgn * f + (d.has_field ? 2ll * sgn * hf[p] : 0ll);
    }
    acc2 = wave_sum_i64(acc2);
    return (int32_t)(acc2 / 2);
}

}  // namespace ptx
