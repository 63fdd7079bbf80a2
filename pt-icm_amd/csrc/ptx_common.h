// ptx_common.h -- definitions shared by the host code and the gfx950 kernels:
// the counter-based random stream, the portable FP64 exp/log the exchange
// decides with, and the per-instance descriptor the kernels read.
//
// The random stream replaces the reference's single mt19937_64
// (include/problem.hh:16-17,70,78).  Every draw is a pure function of
// (seed, instance, purpose, indices), so any launch shape reproduces it.
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PTX_HD __host__ __device__ __forceinline__
#else
#define PTX_HD inline
#endif

namespace ptx {

// Stream tags: low byte of counter word 3, instance id in the upper 24 bits.
enum : uint32_t {
    kTagInit = 1,   // ctr = (i>>2, replica, 0)      word i&3, i = index order
    kTagSweep = 2,  // ctr = (p>>2, replica, sweep)  word p&3, p = colour position
    kTagExch = 3,   // ctr = (a>>2, 0, step)         word a&3, a = attempt
    kTagShuf = 4,   // ctr = (j>>2, half, step)      word j&3, sort key of pair j
    kTagIcm = 5,    // ctr = (t>>2, 0, step)         word t&3, t = ICM slot
    kTagPilot = 0x10
};

struct Philox4 {
    uint32_t v[4];
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
PTX_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    Philox4 r;
    r.v[0] = c0; r.v[1] = c1; r.v[2] = c2; r.v[3] = c3;
    return r;
}

PTX_HD uint32_t ctr3(uint32_t tag, uint32_t inst) { return (tag & 0xFFu) | (inst << 8); }

// Word k&3 of a Philox block without dynamic register indexing.
PTX_HD uint32_t pick(const Philox4& r, uint32_t k) {
    k &= 3u;
    return k == 0 ? r.v[0] : k == 1 ? r.v[1] : k == 2 ? r.v[2] : r.v[3];
}

PTX_HD uint32_t draw_word(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t tag, uint32_t inst,
                          uint32_t word) {
    return pick(philox4x32_10(c0, c1, c2, ctr3(tag, inst), (uint32_t)seed, (uint32_t)(seed >> 32)), word);
}

// ---------------------------------------------------------------------------
// Portable FP64 exp / log.  Only +, -, *, floor and exponent-field bit work,
// evaluated in the written order with contraction off, so the host compiler
// and the gfx950 code generator produce identical bits.
// ---------------------------------------------------------------------------
PTX_HD double bits_as_double(uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __longlong_as_double((long long)b);
#else
    double d;
    memcpy(&d, &b, 8);
    return d;
#endif
}
PTX_HD uint64_t double_as_bits(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint64_t)__double_as_longlong(d);
#else
    uint64_t b;
    memcpy(&b, &d, 8);
    return b;
#endif
}

// e^x: Cody-Waite reduction x = k ln2 + r (fdlibm's split of ln2), degree-14
// Taylor polynomial of e^r for |r| <= 0.35, scaling by 2^k built from bits.
// Results below the normal range flush to zero.
PTX_HD double pexp(double x) {
#pragma clang fp contract(off)
    if (x != x) return x;
    if (x > 709.0) return __builtin_inf();
    if (x < -745.0) return 0.0;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    const double kd = __builtin_floor(x * inv_ln2 + 0.5);
    const double r = (x - kd * ln2_hi) - kd * ln2_lo;
    double p = 1.0 / 87178291200.0;
    p = p * r + 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 1.0 / 2.0;
    p = p * r + 1.0;
    p = p * r + 1.0;
    const int k = (int)kd;
    if (k < -1022) return 0.0;
    if (k > 1023) return __builtin_inf();
    return p * bits_as_double((uint64_t)(k + 1023) << 52);
}

// 2^(-j/32), j = 0..31, as exact binary64 literals.
#define PTX_POW2_NEG_J32                                                                                   \
    {0x1.0000000000000p+0, 0x1.f50765b6e4540p-1, 0x1.ea4afa2a490dap-1, 0x1.dfc97337b9b5fp-1,               \
     0x1.d5818dcfba487p-1, 0x1.cb720dcef9069p-1, 0x1.c199bdd85529cp-1, 0x1.b7f76f2fb5e47p-1,               \
     0x1.ae89f995ad3adp-1, 0x1.a5503b23e255dp-1, 0x1.9c49182a3f090p-1, 0x1.93737b0cdc5e5p-1,               \
     0x1.8ace5422aa0dbp-1, 0x1.82589994cce13p-1, 0x1.7a11473eb0187p-1, 0x1.71f75e8ec5f74p-1,               \
     0x1.6a09e667f3bcdp-1, 0x1.6247eb03a5585p-1, 0x1.5ab07dd485429p-1, 0x1.5342b569d4f82p-1,               \
     0x1.4bfdad5362a27p-1, 0x1.44e086061892dp-1, 0x1.3dea64c123422p-1, 0x1.371a7373aa9cbp-1,               \
     0x1.306fe0a31b715p-1, 0x1.29e9df51fdee1p-1, 0x1.2387a6e756238p-1, 0x1.1d4873168b9aap-1,               \
     0x1.172b83c7d517bp-1, 0x1.11301d0125b51p-1, 0x1.0b5586cf9890fp-1, 0x1.059b0d3158574p-1}

// j for m in [1,2): the largest j in 0..31 with m * tab[j] >= 1, tab[j] =
// 2^(-j/32).  The test is monotone in j (tab decreases, rounding is
// monotone), so from any guess h within one of j three independent tests
// settle it: j = h + 1 if that passes, else h if that passes, else h - 1.
// The guess is floor(32 log2 m): on the device from the single-precision
// hardware logarithm, on the host in double precision.  Either is within
// 1e-4 of 32 log2 m, and j differs from floor(32 log2 m) only where 32
// log2 m is within 1e-13 of an integer, so the guess is within one of j.
// (Three table reads in parallel instead of a bisection's five dependent
// ones, which on the device are five memory round trips.)
PTX_HD int plog_settle(double m, int h, const double* tab) {
#pragma clang fp contract(off)
    const int a = h - 1 < 0 ? 0 : (h - 1 > 31 ? 31 : h - 1);
    const int b = h < 0 ? 0 : (h > 31 ? 31 : h);
    const int c = h + 1 < 0 ? 0 : (h + 1 > 31 ? 31 : h + 1);
    const double tc = tab[c], tb = tab[b];
    if (m * tc >= 1.0) return c;
    if (m * tb >= 1.0) return b;
    return a;
}
PTX_HD int plog_index(double m, const double* tab) {
#if defined(__HIP_DEVICE_COMPILE__)
    return plog_settle(m, (int)(32.0f * __builtin_amdgcn_logf((float)m)), tab);
#else
    return plog_settle(m, (int)(32.0 * __builtin_log2(m)), tab);
#endif
}

// log x: x = 2^k m, m in [1,2); j = largest index with m 2^(-j/32) >= 1
// (plog_index); t = m 2^(-j/32) - 1 in [0, 0.0222); log x = (k + j/32) ln2
// + log1p(t) by its degree-13 Taylor series.  0 -> -inf, negative -> NaN.
PTX_HD double plog(double x) {
#pragma clang fp contract(off)
    const double tab[32] = PTX_POW2_NEG_J32;
    if (x != x || x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    int k = 0;
    uint64_t b = double_as_bits(x);
    if ((b >> 52) == 0) {
        x = x * 0x1p54;
        k = -54;
        b = double_as_bits(x);
    }
    k += (int)((b >> 52) & 0x7FF) - 1023;
    const double m = bits_as_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    const int lo = plog_index(m, tab);
    const double t = m * tab[lo] - 1.0;
    double p = 1.0 / 13.0;
    p = p * t + -1.0 / 12.0;
    p = p * t + 1.0 / 11.0;
    p = p * t + -1.0 / 10.0;
    p = p * t + 1.0 / 9.0;
    p = p * t + -1.0 / 8.0;
    p = p * t + 1.0 / 7.0;
    p = p * t + -1.0 / 6.0;
    p = p * t + 1.0 / 5.0;
    p = p * t + -1.0 / 4.0;
    p = p * t + 1.0 / 3.0;
    p = p * t + -1.0 / 2.0;
    p = p * t + 1.0;
    p = p * t;
    const double ln2 = 6.93147180559945286227e-01;
    const double kj = (double)k + (double)lo * 0.03125;
    return kj * ln2 + p;
}

// log(u) for u = r * 2^-32 (the exchange's uniform draw); -inf for r = 0.
PTX_HD double uniform_log(uint32_t r) {
#pragma clang fp contract(off)
    if (r == 0) return -__builtin_inf();
    return plog((double)r * 0x1p-32);
}

// ---------------------------------------------------------------------------
// Device data layout.  Positions p are the colour-sorted vertex order; all
// device arrays are indexed by position.  The adjacency of position p is a
// row of records padded to 16 bytes (one 128-bit load reads 8 16-bit or 4
// 32-bit records); rowinfo[p] = (row start in 16-byte units) << 9 | degree.
// ---------------------------------------------------------------------------
struct InstDesc {
    int32_t n;         // spins
    int32_t nw;        // 32-bit words per replica state
    int32_t nt;        // temperatures per chain
    int32_t nrep;      // chains * nt (= slots)
    int32_t ncls;      // colour classes
    int32_t cls0;      // positions of colour class 0
    int32_t fmax;      // threshold-table width (max |dE|/2)
    int32_t nc;        // ICM slots
    int32_t has_field; // any h != 0
    int32_t neg;       // any coupling < 0 (0: unit-format records carry no sign, the gathers skip it)
    int32_t kbit;      // unit format: sign bit of a record (15, or the bit of the split sweep's sign table)
    int64_t pos_off;     // -> rowinfo[n], h[n], pos2orig[n]
    int64_t adj_off16;   // -> first 16-byte unit of the instance's adjacency rows
    int64_t adj_len16;   // 16-byte units of adjacency rows
    int64_t cls_off;     // -> class_start[ncls+1]
    int64_t thr_off;     // -> thr[nt][fmax]
    int64_t beta_off;    // -> beta[nrep] (slot betas, FP64)
    int64_t rep_off;     // -> per-replica arrays (E, best, flips, rep2slot, slot2rep, swap stats)
    int64_t spin_off;    // -> spins[nrep][nw]
    int64_t hist_off;    // -> hist[nrep][hist_bins]
    int32_t hist_bins;
    int32_t e_base;      // minimum possible energy (histogram bin 0)
    uint32_t inst_id;    // random-stream instance id
    int32_t max_deg;
    int32_t deg_mean;    // adjacency entries per spin, rounded down
    int32_t lat_dim;     // 2 or 3: periodic lattice (bit-set cluster growth), else 0
    int32_t lat_L;
    int32_t lat_R;       // 32-bit words per lattice row (rows padded to whole words)
    int32_t lat_words;   // words of a lattice-order bit set: lat_R * L^(lat_dim-1)
    int64_t lat_off;     // -> lat_o2p[n] (index -> position) and lat_obit[n] (position -> word<<5|bit)
    int64_t jn_off;      // -> edge signs [lat_dim][rows] (bit x of row r: the coupling of the site to its +axis
                         //    neighbour is negative) for cut energies on bit sets, or -1
    // Multispin sweep of the split kernels: 2 or 3 for a periodic L^D lattice
    // with even L, couplings +-1 and no field, whose two colour classes are
    // the site parities with positions in index order (class c, row r =
    // y + L z, site x: position class_start[c] + r L/2 + x/2); else 0.
    int32_t lsw_dim;
    int32_t lsw_L;
    int64_t lsw_off;     // -> per class and window of 32 class positions, kLswGroup words: the edge
                         //    signs to the +x, -x, +y, -y, +z, -z neighbours (bit = coupling < 0), 0, 0,
                         //    and the shape masks par, k = 0, k = L/2 - 1, y = 0, y = L-1, z = 0, z = L-1, 0
    // Replica-sliced sweep of the split kernels (k_sweep_rs, 32 replicas per
    // wave): 1 for unit records, no field, degree <= 63 and no multispin
    // lattice sweep, when the engine allows it; else 0
    int32_t rs;
};

constexpr int kLswGroup = 16;  // words per (class, window) of the multispin sweep's table

}  // namespace ptx
