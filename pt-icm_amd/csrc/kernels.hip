// kernels.hip -- the gfx950 kernels of one PT step (split form) plus state
// initialisation and device self-tests.
//
//   k_init      random Bernoulli(1/2) states (initialization.cc:63-76), energies
//   k_sweep     Metropolis sweeps, one wavefront per replica (replica.cc:5-38)
//   k_exchange  histogram of slot energies + (NN)_a exchange as a temperature
//               permutation, one wavefront per instance (replica_exchange.cc:5-100,
//               cae_estimator.cc:122-139 in pilot mode)
//   k_icm       Houdayer cluster move, one wavefront per slot pair (icm.cc:8-215)
#include "kernels.hpp"

namespace ptx {

// ---------------------------------------------------------------------------
// k_init: one wave per replica.  mode 0 draws the initial state; mode 1 keeps
// the stored state and only recomputes its energy.
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(W* kWave) void k_init(DevArgs a, int mode) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int2 it = a.items[blockIdx.x];
    const InstDesc d = a.desc[it.x];
    const int r = it.y + wave;
    if (r >= d.nrep) return;
    uint8_t* sp = smem + wave * a.spin_stride;
    const int64_t gr = d.rep_off + r;
    uint32_t* words = a.spins + d.spin_off + (int64_t)r * d.nw;
    if (mode == 0) {
        const uint32_t c3 = ctr3(kTagInit | a.tag_base, d.inst_id);
        for (int p = lane; p < d.nw * 32; p += kWave) {
            uint8_t b = 0;
            if (p < d.n) {
                const uint32_t i = (uint32_t)a.p2o[d.p2o_off + p];
                const Philox4 x = philox4x32_10(i >> 2, (uint32_t)r, 0u, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
                b = (uint8_t)(pick(x, i) >> 31);
            }
            sp[p] = b;
        }
        wave_sync();
        pack_bits(sp, d.nw, words, lane);
        pack_bits(sp, d.nw, a.best_spins + d.spin_off + (int64_t)r * d.nw, lane);
    } else {
        expand_bits(words, d.nw, sp, lane);
        wave_sync();
    }
    const int32_t E = energy_bytes(a, d, sp, lane);
    if (lane == 0) {
        a.E[gr] = E;
        if (mode == 0) {
            a.bestE[gr] = 0x7FFFFFFF;
            a.best_sweep[gr] = 0;
            a.flips[gr] = 0;
            a.slot2rep[gr] = r;
            a.rep2slot[gr] = r;
            a.swap_att[gr] = 0;
            a.swap_succ[gr] = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// k_sweep: W wavefronts per workgroup, each sweeping one replica of the
// workgroup's instance.  LDS per wave: spin bytes + the replica's threshold row.
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(W* kWave) void k_sweep(DevArgs a, uint32_t step) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int2 it = a.items[blockIdx.x];
    const InstDesc d = a.desc[it.x];
    const int r = it.y + wave;
    if (r >= d.nrep) return;
    uint8_t* sp = smem + wave * a.spin_stride;
    uint32_t* thr_row = reinterpret_cast<uint32_t*>(smem + W * a.spin_stride) + wave * a.thr_stride;
    const int64_t gr = d.rep_off + r;
    uint32_t* words = a.spins + d.spin_off + (int64_t)r * d.nw;
    uint32_t* best_words = a.best_spins + d.spin_off + (int64_t)r * d.nw;
    const int t = a.rep2slot[gr] % d.nt;
    for (int m = lane; m < d.fmax; m += kWave) thr_row[m] = a.thr[d.thr_off + (int64_t)t * d.fmax + m];
    expand_bits(words, d.nw, sp, lane);
    SweepState st;
    st.E = a.E[gr];
    st.bestE = a.bestE[gr];
    st.best_sweep = a.best_sweep[gr];
    st.flips = 0;
    wave_sync();
    for (int sw = 0; sw < a.sweeps; ++sw) {
        const uint32_t gsweep = step * (uint32_t)a.sweeps + (uint32_t)sw;
        sweep_replica(a, d, r, gsweep, sp, thr_row, best_words, st, lane);
    }
    pack_bits(sp, d.nw, words, lane);
    if (lane == 0) {
        a.E[gr] = st.E;
        a.bestE[gr] = st.bestE;
        a.best_sweep[gr] = st.best_sweep;
        a.flips[gr] += st.flips;
    }
}

// ---------------------------------------------------------------------------
// k_exchange: one wavefront per instance.  Lane-parallel: histogram update,
// pair orders (random sort keys), log-uniforms of every attempt.  Lane 0 then
// walks the attempts in order; each reads the permutation the previous ones
// left (replica_exchange.cc:11-14,74).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void k_exchange(DevArgs a, uint32_t step, int nrep_max) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x;
    const InstDesc d = a.desc[blockIdx.x];
    const int nrep = d.nrep, nt = d.nt, np = nt - 1;
    // LDS: 40 bytes per slot (2(nt-1) attempts and two orders fit in nrep slots)
    double* beta = reinterpret_cast<double*>(smem);
    double* logu = beta + nrep_max;
    int32_t* E = reinterpret_cast<int32_t*>(logu + nrep_max);
    int32_t* s2r = E + nrep_max;
    uint32_t* att = reinterpret_cast<uint32_t*>(s2r + nrep_max);
    uint32_t* succ = att + nrep_max;
    uint32_t* key = succ + nrep_max;
    int32_t* perm = reinterpret_cast<int32_t*>(key + nrep_max);  // [2][np]
    const int64_t ro = d.rep_off;
    for (int s = lane; s < nrep; s += kWave) {
        beta[s] = a.beta[d.beta_off + s];
        s2r[s] = a.slot2rep[ro + s];
        E[s] = a.E[ro + s];
        att[s] = 0;
        succ[s] = 0;
    }
    __syncthreads();
    if (a.hist_on && (int64_t)step >= a.hist_start) {
        for (int s = lane; s < nrep; s += kWave) {
            const int e = E[s2r[s]];
            const int bin = (e - d.e_base) >> 1;
            atomicAdd(&a.hist[d.hist_off + (int64_t)s * d.hist_bins + bin], 1u);
        }
    }
    if (a.no_exchange || np <= 0) return;
    const int nhalf = a.pilot ? 1 : 2;
    const int nattempt = np * nhalf;
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const uint32_t c3s = ctr3(kTagShuf | a.tag_base, d.inst_id);
    for (int h = 0; h < nhalf; ++h) {
        for (int j = lane; j < np; j += kWave) {
            const Philox4 x = philox4x32_10((uint32_t)j >> 2, (uint32_t)h, step, c3s, k0, k1);
            key[j] = pick(x, (uint32_t)j);
        }
        __syncthreads();
        for (int j = lane; j < np; j += kWave) {
            const uint32_t kj = key[j];
            int rank = 0;
            for (int i = 0; i < np; ++i) {
                const uint32_t ki = key[i];
                rank += (ki < kj) || (ki == kj && i < j);
            }
            perm[h * np + rank] = j;
        }
        __syncthreads();
    }
    const uint32_t c3e = ctr3(kTagExch | a.tag_base, d.inst_id);
    for (int t = lane; t < nattempt; t += kWave) {
        const Philox4 x = philox4x32_10((uint32_t)t >> 2, 0u, step, c3e, k0, k1);
        logu[t] = uniform_log(pick(x, (uint32_t)t));
    }
    __syncthreads();
    if (lane == 0) {
        unsigned long long nsw = 0;
        for (int t = 0; t < nattempt; ++t) {
            const int chain = a.pilot ? 0 : (t & 1);
            const int k = perm[(t / np) * np + (t % np)] + chain * nt;
            att[k] += 1u;
            const int r1 = s2r[k], r2 = s2r[k + 1];
            const double lb = (beta[k] - beta[k + 1]) * (double)(E[r1] - E[r2]);
            if (logu[t] < lb) {
                s2r[k] = r2;
                s2r[k + 1] = r1;
                succ[k] += 1u;
                ++nsw;
            }
        }
        a.swaps[blockIdx.x] += nsw;
    }
    __syncthreads();
    for (int s = lane; s < nrep; s += kWave) {
        a.slot2rep[ro + s] = s2r[s];
        a.rep2slot[ro + s2r[s]] = s;
        a.swap_att[ro + s] += att[s];
        a.swap_succ[ro + s] += succ[s];
    }
}

// ---------------------------------------------------------------------------
// k_icm: one wavefront per (instance, slot t) pair, t in [nt-nc, nt).
// LDS: both states, the disagreement set D, the visited set V (bit words)
// and a queue of up to n/2 positions.  The cluster is grown breadth-first
// from a uniform seed in D; the reference grows it depth-first
// (icm.cc:37-72), which gives the same component.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void k_icm(DevArgs a, uint32_t step, int nw_max) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x;
    const int2 pr = a.pairs[blockIdx.x];
    const InstDesc d = a.desc[pr.x];
    const int t = pr.y;
    const int n = d.n, nw = d.nw;
    uint32_t* S1 = reinterpret_cast<uint32_t*>(smem);
    uint32_t* S2 = S1 + nw_max;
    uint32_t* D = S2 + nw_max;
    uint32_t* V = D + nw_max;
    int32_t* tail = reinterpret_cast<int32_t*>(V + nw_max);
    uint16_t* Q = reinterpret_cast<uint16_t*>(tail + 4);
    const int64_t ro = d.rep_off;
    const int r1 = a.slot2rep[ro + t], r2 = a.slot2rep[ro + t + d.nt];
    uint32_t* g1 = a.spins + d.spin_off + (int64_t)r1 * nw;
    uint32_t* g2 = a.spins + d.spin_off + (int64_t)r2 * nw;
    const uint32_t lastmask = (n & 31) ? ((1u << (n & 31)) - 1u) : 0xFFFFFFFFu;
    int cnt = 0;
    for (int w = lane; w < nw; w += kWave) {
        const uint32_t m = (w == nw - 1) ? lastmask : 0xFFFFFFFFu;
        const uint32_t x1 = g1[w] & m, x2 = g2[w] & m;
        S1[w] = x1;
        S2[w] = x2;
        D[w] = x1 ^ x2;
        V[w] = 0u;
        cnt += __popc(x1 ^ x2);
    }
    cnt = wave_sum_i32(cnt);
    const int32_t* __restrict__ rp = a.rowptr + d.rowptr_off;
    const int32_t* __restrict__ hf = a.hfield + d.h_off;
    long long dE1 = 0, dE2 = 0;
    __syncthreads();
    if (cnt > n / 2) {  // icm.cc:20-28: flip replica 1 globally
        if (d.has_field) {
            for (int p = lane; p < n; p += kWave) {
                const int s1 = (S1[p >> 5] >> (p & 31)) & 1u;
                dE1 += -2ll * (s1 ? 1 : -1) * hf[p];
            }
        }
        __syncthreads();
        for (int w = lane; w < nw; w += kWave) {
            const uint32_t m = (w == nw - 1) ? lastmask : 0xFFFFFFFFu;
            S1[w] = ~S1[w] & m;
            D[w] = ~D[w] & m;
        }
        cnt = n - cnt;
        __syncthreads();
    }
    if (cnt > 0) {
        const Philox4 x = philox4x32_10((uint32_t)t >> 2, 0u, step, ctr3(kTagIcm | a.tag_base, d.inst_id),
                                        (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
        const int k = (int)(((unsigned long long)pick(x, (uint32_t)t) * (unsigned long long)cnt) >> 32);
        // k-th element of D in position order: per-64-word chunks, inclusive scan
        int seed_pos = -1, base = 0;
        for (int w0 = 0; w0 < nw && seed_pos < 0; w0 += kWave) {
            const int w = w0 + lane;
            const uint32_t dw = w < nw ? D[w] : 0u;
            const int c = __popc(dw);
            int incl = c;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const int y = __shfl_up(incl, o, kWave);
                if (lane >= o) incl += y;
            }
            const int total = __shfl(incl, kWave - 1, kWave);
            if (k < base + total) {
                const bool mine = (k < base + incl) && (k >= base + incl - c);
                const unsigned long long ball = __ballot(mine);
                const int owner = __ffsll((long long)ball) - 1;
                int sel = -1;
                if (lane == owner) {
                    int kk = k - (base + incl - c);
                    uint32_t v = dw;
                    for (int b = 0; b < 32; ++b) {
                        if ((v >> b) & 1u) {
                            if (kk == 0) { sel = 32 * w + b; break; }
                            --kk;
                        }
                    }
                }
                seed_pos = __shfl(sel, owner, kWave);
            }
            base += total;
        }
        if (lane == 0) {
            V[seed_pos >> 5] |= 1u << (seed_pos & 31);
            Q[0] = (uint16_t)seed_pos;
            tail[0] = 1;
        }
        __syncthreads();
        int head = 0;
        while (true) {
            const int tl = tail[0];
            __syncthreads();
            if (head >= tl) break;
            const int idx = head + lane;
            if (idx < tl) {
                const int p = Q[idx];
                for (int q = rp[p]; q < rp[p + 1]; ++q) {
                    const int j = (int)(a.adj[q] & 0xFFFFu);
                    const uint32_t bit = 1u << (j & 31);
                    if ((D[j >> 5] & bit) && !(V[j >> 5] & bit)) {
                        const uint32_t old = atomicOr(&V[j >> 5], bit);
                        if (!(old & bit)) {
                            const int pos = atomicAdd(&tail[0], 1);
                            Q[pos] = (uint16_t)j;
                        }
                    }
                }
            }
            head = (tl - head > kWave) ? head + kWave : tl;
            __syncthreads();
        }
        const int csize = tail[0];
        // energy change of flipping the cluster C in each replica: boundary
        // edges (p in C, j not in C) and the field terms of C
        for (int idx = lane; idx < csize; idx += kWave) {
            const int p = Q[idx];
            const int s1 = ((S1[p >> 5] >> (p & 31)) & 1u) ? 1 : -1;
            const int s2 = ((S2[p >> 5] >> (p & 31)) & 1u) ? 1 : -1;
            long long b1 = 0, b2 = 0;
            for (int q = rp[p]; q < rp[p + 1]; ++q) {
                const uint32_t rec = a.adj[q];
                const int j = (int)(rec & 0xFFFFu);
                if ((V[j >> 5] >> (j & 31)) & 1u) continue;
                const int w = (int32_t)rec >> 16;
                b1 += (long long)w * (((S1[j >> 5] >> (j & 31)) & 1u) ? 1 : -1);
                b2 += (long long)w * (((S2[j >> 5] >> (j & 31)) & 1u) ? 1 : -1);
            }
            dE1 += -2ll * s1 * b1;
            dE2 += -2ll * s2 * b2;
            if (d.has_field) {
                dE1 += -2ll * s1 * hf[p];
                dE2 += -2ll * s2 * hf[p];
            }
        }
        __syncthreads();
        for (int w = lane; w < nw; w += kWave) {
            S1[w] ^= V[w];
            S2[w] ^= V[w];
        }
        __syncthreads();
    }
    dE1 = wave_sum_i64(dE1);
    dE2 = wave_sum_i64(dE2);
    for (int w = lane; w < nw; w += kWave) {
        g1[w] = S1[w];
        g2[w] = S2[w];
    }
    if (lane == 0) {
        a.E[ro + r1] += (int32_t)dE1;
        a.E[ro + r2] += (int32_t)dE2;
    }
}

// ---------------------------------------------------------------------------
// Self-tests.
// ---------------------------------------------------------------------------
__global__ void k_selftest_fp64(const double* x, int64_t n, double* eo, double* lo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        eo[i] = pexp(x[i]);
        lo[i] = plog(x[i]);
    }
}

__global__ void k_selftest_philox(const uint32_t* ctr, int64_t n, uint32_t k0, uint32_t k1, uint32_t* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const Philox4 r = philox4x32_10(ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3], k0, k1);
        out[4 * i] = r.v[0];
        out[4 * i + 1] = r.v[1];
        out[4 * i + 2] = r.v[2];
        out[4 * i + 3] = r.v[3];
    }
}

// explicit instantiations used by the host launcher
template __global__ void k_init<1>(DevArgs, int);
template __global__ void k_init<2>(DevArgs, int);
template __global__ void k_init<4>(DevArgs, int);
template __global__ void k_init<8>(DevArgs, int);
template __global__ void k_sweep<1>(DevArgs, uint32_t);
template __global__ void k_sweep<2>(DevArgs, uint32_t);
template __global__ void k_sweep<4>(DevArgs, uint32_t);
template __global__ void k_sweep<8>(DevArgs, uint32_t);

}  // namespace ptx
