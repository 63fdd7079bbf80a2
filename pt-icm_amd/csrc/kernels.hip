// kernels.hip -- the gfx950 kernels.
//
// Split form (one launch per phase of a PT step; any instance size):
//   k_init      Bernoulli(1/2) initial states (initialization.cc:63-76) and energies
//   k_sweep     Metropolis sweeps, one wavefront per replica (replica.cc:5-38)
//   k_exchange  slot-energy histogram + (NN)_a exchange, one wavefront per
//               instance (replica_exchange.cc:5-100; cae_estimator.cc:122-139
//               in pilot mode)
//   k_icm       Houdayer cluster move, 1, 2 or 4 wavefronts per slot pair (icm.cc:8-215)
// Fused form (batched instances): k_fused runs whole PT steps -- sweeps of
// every replica, the exchange and the cluster moves -- for one instance per
// workgroup with the adjacency staged in LDS, several steps per launch.
#include "kernels.hpp"

namespace ptx {

// ---------------------------------------------------------------------------
// k_init: one wave per replica.  mode 0 draws the initial state; mode 1 keeps
// the stored state and only recomputes its energy.
// ---------------------------------------------------------------------------
template <int Fmt, int W>
__global__ __launch_bounds__(W* kWave) void k_init(DevArgs a, int mode) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int2 it = a.items[blockIdx.x];
    if (it.x < 0) return;  // idle item padding an XCD's list
    const InstDesc d = a.desc[it.x];
    const int r = it.y + wave;
    if (r >= d.nrep) return;
    uint8_t* sp = smem + wave * a.spin_stride;
    const int64_t gr = d.rep_off + r;
    uint32_t* words = a.spins + d.spin_off + (int64_t)r * d.nw;
    if (mode == 0) {
        const uint32_t c3 = ctr3(kTagInit | a.tag_base, d.inst_id);
        for (int p = lane; p < spin_bytes(d.nw); p += kWave) {  // includes the zero sentinel word
            uint8_t b = 0;
            if (p < d.n) {
                const uint32_t i = (uint32_t)a.p2o[d.pos_off + p];
                const Philox4 x =
                    philox4x32_10(i >> 2, (uint32_t)r, 0u, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
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
    const Csr<Fmt> csr{a.rowinfo + d.pos_off, a.adj + d.adj_off16 * 16, d.neg, d.kbit, 0,
                       sentinel_pos(d.nw)};
    const int32_t E = energy_bytes(csr, d.n, a.hfield + d.pos_off, d.has_field != 0, sp, lane);
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
// workgroup's instance.  LDS per wave: spin bytes (lattice instances with the
// multispin sweep: state bits and uphill level masks) + the replica's
// threshold row.
// ---------------------------------------------------------------------------
template <int Fmt, int W>
__global__ __launch_bounds__(W* kWave) void k_sweep(DevArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int2 it = a.items[blockIdx.x];
    if (it.x < 0) return;  // idle item padding an XCD's list
    const uint32_t step = a.inst_step[it.x];
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
    SweepState st;
    st.E = a.E[gr];
    st.bestE = a.bestE[gr];
    st.best_sweep = a.best_sweep[gr];
    st.flips = 0;
    const uint32_t c3 = ctr3(kTagSweep | a.tag_base, d.inst_id);
    if (d.lsw_dim > 0) {  // periodic lattice: multispin sweep on the state bits
        uint32_t* S0 = reinterpret_cast<uint32_t*>(sp);
        uint32_t* S = S0 + 1;                      // state words; S[-1], S[nw], S[nw + 1] are zero
        uint32_t* Z = S0 + lsw_state_words(d.nw);  // level masks, zero pads included
        for (int w = lane; w < lsw_state_words(d.nw); w += kWave) S0[w] = w >= 1 && w <= d.nw ? words[w - 1] : 0u;
        for (int w = lane; w < d.lsw_dim * lsw_zwords(d.n); w += kWave) Z[w] = 0u;
        wave_sync();
        const uint32_t* tab = a.lsign + d.lsw_off;
        const int32_t* cls = a.cls + d.cls_off;
        for (int sw = 0; sw < a.sweeps; ++sw) {
            const uint32_t gsweep = step * (uint32_t)a.sweeps + (uint32_t)sw;
            if (d.lsw_dim == 3)
                sweep_lattice<3>(tab, d.lsw_L, d.n, cls, d.nw, c3, a.seed, r, gsweep, S, Z, thr_row, best_words, st,
                                 lane);
            else
                sweep_lattice<2>(tab, d.lsw_L, d.n, cls, d.nw, c3, a.seed, r, gsweep, S, Z, thr_row, best_words, st,
                                 lane);
        }
        for (int w = lane; w < d.nw; w += kWave) words[w] = S[w];
    } else {
        expand_bits(words, d.nw, sp, lane);
        wave_sync();
        if (a.tab) {  // the sign table's negated bytes at 2^kbit
            const uint32_t* src = reinterpret_cast<const uint32_t*>(sp);
            uint32_t* dst = reinterpret_cast<uint32_t*>(sp + (1 << d.kbit));
            for (int q = lane; q < spin_bytes(d.nw) / 4; q += kWave) dst[q] = src[q] ^ 0x01010101u;
            wave_sync();
        }
        const Csr<Fmt> csr{a.rowinfo + d.pos_off, a.adj + d.adj_off16 * 16, d.neg, d.kbit, a.tab,
                           sentinel_pos(d.nw)};
        // rows of one chunk each (lattices, sparse graphs): fixed-stride rows, no chunk loop
        const bool one = d.max_deg <= Csr<Fmt>::kPerChunk;
        for (int sw = 0; sw < a.sweeps; ++sw) {
            const uint32_t gsweep = step * (uint32_t)a.sweeps + (uint32_t)sw;
            if (one)
                sweep_replica<Fmt, true>(csr, a.cls + d.cls_off, d.ncls, a.hfield + d.pos_off, d.has_field != 0, d.nw, c3,
                                         a.seed, r, gsweep, sp, thr_row, best_words, st, lane);
            else
                sweep_replica<Fmt, false>(csr, a.cls + d.cls_off, d.ncls, a.hfield + d.pos_off, d.has_field != 0, d.nw,
                                          c3, a.seed, r, gsweep, sp, thr_row, best_words, st, lane);
        }
        pack_bits(sp, d.nw, words, lane);
    }
    if (lane == 0) {
        a.E[gr] = st.E;
        a.bestE[gr] = st.bestE;
        a.best_sweep[gr] = st.best_sweep;
        a.flips[gr] += st.flips;
    }
}

// ---------------------------------------------------------------------------
// k_sweep_rs: W waves per 32 replicas (item (instance, r0)) of an instance
// swept replica-sliced (InstDesc::rs).  LDS: the per-class sums of the
// waves, then X[32 (nw + 1)], the position words of the 32 replicas (the
// last 32 the zero block the padding records read), then the replicas'
// threshold rows [32][fmax].
// ---------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(W* kWave) __attribute__((amdgpu_waves_per_eu(4))) void k_sweep_rs(DevArgs a) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int tid = threadIdx.x, lane = tid % kWave, wave = tid / kWave;
    const int2 it = a.items[blockIdx.x];
    if (it.x < 0) return;  // idle item padding an XCD's list
    const uint32_t step = a.inst_step[it.x];
    const InstDesc d = a.desc[it.x];
    const int r0 = it.y, nb = min(32, d.nrep - r0), nw = d.nw;
    SlicedShared& sh = *reinterpret_cast<SlicedShared*>(smem);
    uint32_t* X = reinterpret_cast<uint32_t*>(smem + kSlicedShared);
    uint32_t* T = X + 32 * (nw + 1);
    for (int j = wave; j < nb; j += W) {
        const int t = a.rep2slot[d.rep_off + r0 + j] % d.nt;
        for (int m = lane; m < d.fmax; m += kWave) T[j * d.fmax + m] = a.thr[d.thr_off + (int64_t)t * d.fmax + m];
    }
    // replica-major state words -> position words: half-wave hw transposes
    // block w = hw, hw + 2 W, ... (lane j brings replica r0 + j's word w);
    // block nw is the sentinel's, zero
    uint32_t* words = a.spins + d.spin_off + (int64_t)r0 * nw;
    const int l32 = tid & 31;
    for (int w = tid >> 5; w <= nw; w += 2 * W) {
        const uint32_t v = (w < nw && l32 < nb) ? words[(int64_t)l32 * nw + w] : 0u;
        X[32 * w + l32] = transpose_lanes32(v, l32);
    }
    const int64_t gr = d.rep_off + r0 + lane;
    SweepState st{0, 0, 0, 0ull};
    if (wave == 0 && lane < nb) {
        st.E = a.E[gr];
        st.bestE = a.bestE[gr];
        st.best_sweep = a.best_sweep[gr];
    }
    __syncthreads();
    const uint32_t c3 = ctr3(kTagSweep | a.tag_base, d.inst_id);
    const uint32_t k0 = (uint32_t)a.seed, k1 = (uint32_t)(a.seed >> 32);
    const uint4* adj = reinterpret_cast<const uint4*>(a.adj + d.adj_off16 * 16);
    uint32_t* best_words = a.best_spins + d.spin_off + (int64_t)r0 * nw;
    for (int sw = 0; sw < a.sweeps; ++sw) {
        const uint32_t gsweep = step * (uint32_t)a.sweeps + (uint32_t)sw;
        sweep_sliced<W>(d, a.rowinfo + d.pos_off, adj, a.cls + d.cls_off, c3, k0, k1, r0, nb, gsweep, X, T, sh,
                        best_words, st, tid);
    }
    for (int w = tid >> 5; w < nw; w += 2 * W) {
        const uint32_t v = transpose_lanes32(X[32 * w + l32], l32);
        if (l32 < nb) words[(int64_t)l32 * nw + w] = v;
    }
    if (wave == 0 && lane < nb) {
        a.E[gr] = st.E;
        a.bestE[gr] = st.bestE;
        a.best_sweep[gr] = st.best_sweep;
        a.flips[gr] += st.flips;
    }
}

// Histogram of the slot energies after the sweeps (the point where
// gather_samples logs them, cae_estimator.cc:117-121): energy = e_base + 2*bin.
__device__ __forceinline__ void record_histogram(const DevArgs& a, const InstDesc& d, uint32_t step, const int32_t* E,
                                                 const int32_t* s2r, int lane) {
    if (!a.hist_on || (int64_t)step < a.hist_start) return;
    for (int s = lane; s < d.nrep; s += kWave) {
        const int bin = (E[s2r[s]] - d.e_base) >> 1;
        // energies lie in [e_base, -e_base]; a sample outside the table is an
        // engine fault: it is counted (fault word 1, host-mapped) instead of written
        if (bin >= 0 && bin < d.hist_bins) atomicAdd(&a.hist[d.hist_off + (int64_t)s * d.hist_bins + bin], 1u);
        else __hip_atomic_fetch_add(a.fault + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// ---------------------------------------------------------------------------
// k_exchange: one wavefront per instance (LDS: 48 bytes per slot).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWave) void k_exchange(DevArgs a, int nrep_max) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int lane = threadIdx.x;
    const int inst = a.order[blockIdx.x];  // the active instances
    const uint32_t step = a.inst_step[inst];
    const InstDesc d = a.desc[inst];
    const int nrep = d.nrep;
    double* beta = reinterpret_cast<double*>(smem);
    ExchScratch x;
    x.logu = beta + nrep_max;
    int32_t* E = reinterpret_cast<int32_t*>(x.logu + nrep_max);
    int32_t* s2r = E + nrep_max;
    uint32_t* att = reinterpret_cast<uint32_t*>(s2r + nrep_max);
    uint32_t* succ = att + nrep_max;
    x.key = succ + nrep_max;
    x.perm = reinterpret_cast<int32_t*>(x.key + nrep_max);
    x.tpos = x.perm + nrep_max;
    x.level = x.tpos + nrep_max;
    const int64_t ro = d.rep_off;
    for (int s = lane; s < nrep; s += kWave) {
        beta[s] = a.beta[d.beta_off + s];
        s2r[s] = a.slot2rep[ro + s];
        E[s] = a.E[ro + s];
        att[s] = 0;
        succ[s] = 0;
    }
    wave_sync();
    record_histogram(a, d, step, E, s2r, lane);
    if (a.no_exchange) return;
    const unsigned long long nsw = exchange_wave(d.nt, nrep, a.pilot != 0, a.seed, a.tag_base, d.inst_id, step, beta,
                                                 E, s2r, att, succ, x, lane);
    const unsigned long long total = (unsigned long long)wave_sum_i64((long long)nsw);
    if (lane == 0) a.swaps[inst] += total;
    wave_sync();
    for (int s = lane; s < nrep; s += kWave) {
        a.slot2rep[ro + s] = s2r[s];
        a.rep2slot[ro + s2r[s]] = s;
        a.swap_att[ro + s] += att[s];
        a.swap_succ[ro + s] += succ[s];
    }
}

// ---------------------------------------------------------------------------
// k_icm: one workgroup of W waves per (instance, slot t) pair, t in [nt-nc, nt).
// LDS: the reduction slots, the ICM scratch of the largest instance (S1, S2,
// D, V, tail, a queue of 32 * nw_max + 16 bytes) and two lattice-order bit
// sets.  One wave grows lattice clusters as bit sets.
// ---------------------------------------------------------------------------
template <int Fmt, int W>
__global__ __launch_bounds__(W* kWave) void k_icm(DevArgs a, int nw_max, int lat_words_max) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int tid = threadIdx.x;
    const int2 pr = a.pairs[blockIdx.x];
    if (pr.x < 0) return;  // idle item padding an XCD's list
    const uint32_t step = a.inst_step[pr.x];
    const InstDesc d = a.desc[pr.x];
    const int t = pr.y;
    IcmScratch sc;
    sc.red = reinterpret_cast<long long*>(smem);
    sc.S1 = reinterpret_cast<uint32_t*>(smem + 8 * kIcmRedSlots);
    sc.S2 = sc.S1 + (nw_max + 1);
    sc.D = sc.S2 + (nw_max + 1);
    sc.V = sc.D + (nw_max + 1);
    sc.tail = reinterpret_cast<int32_t*>(sc.V + (nw_max + 1));
    sc.Q = reinterpret_cast<uint16_t*>(sc.tail + 4);
    sc.Do = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(sc.Q) + 32 * nw_max + 16);
    sc.Vo = sc.Do + lat_words_max;
    const int64_t ro = d.rep_off;
    const int r1 = a.slot2rep[ro + t], r2 = a.slot2rep[ro + t + d.nt];
    uint32_t* g1 = a.spins + d.spin_off + (int64_t)r1 * d.nw;
    uint32_t* g2 = a.spins + d.spin_off + (int64_t)r2 * d.nw;
    const Csr<Fmt> csr{a.rowinfo + d.pos_off, a.adj + d.adj_off16 * 16, d.neg, d.kbit, 0,
                       sentinel_pos(d.nw)};
    long long dE1 = 0, dE2 = 0;
    const uint32_t c3i = ctr3(kTagIcm | a.tag_base, d.inst_id);
    if (W > 1 && d.lat_dim > 0) {  // lattices grow as bit sets on the first wave
        if (tid >= kWave) return;
        icm_team<Fmt, false>(csr, a.hfield + d.pos_off, d.has_field != 0, d.n, d.nw, t, step, a.seed, c3i, g1,
                             g2, sc, dE1, dE2, tid, kWave, true, lat_view(a, d));
    } else {
        icm_team<Fmt, (W > 1)>(csr, a.hfield + d.pos_off, d.has_field != 0, d.n, d.nw, t, step, a.seed, c3i,
                               g1, g2, sc, dE1, dE2, tid, W * kWave, true, W == 1 ? lat_view(a, d) : LatView{},
                               true, a.E[ro + r1], a.E[ro + r2]);
    }
    if (tid == 0) {
        a.E[ro + r1] += (int32_t)dE1;
        a.E[ro + r2] += (int32_t)dE2;
    }
}

// ---------------------------------------------------------------------------
// Fused per-instance kernel: one workgroup of W waves owns one instance for
// nsteps PT steps.  LDS holds the adjacency (16- or 32-bit records), the
// threshold table, betas, energies, the permutation, the exchange scratch and
// one scratch area per wave (spin bytes during sweeps, ICM sets during
// cluster moves).  Spin words stay in global memory (L2-resident).
// ---------------------------------------------------------------------------
struct FusedLayout {
    int info, adj, thr, beta, E, s2r, r2s, att, succ, logu, key, perm, tpos, level, scratch, scratch_stride, total;
};

__host__ __device__ inline int align16(int x) { return (x + 15) & ~15; }

// csr_lds = 0: the adjacency stays in global memory (read through L2), only
// the per-replica state is staged.
// icm = 0: no cluster moves, so the per-wave scratch only holds spin bytes.
__host__ __device__ inline FusedLayout fused_layout(int n, int adj_len16, int nrep, int nt, int fmax, int W,
                                                    int csr_lds, int icm) {
    FusedLayout L;
    int o = 0;
    L.info = o;
    o += csr_lds ? align16(4 * n) : 16;
    L.adj = o;
    o += csr_lds ? 16 * (adj_len16 > 0 ? adj_len16 : 1) : 16;
    L.thr = o;
    o += align16(4 * nt * (fmax > 0 ? fmax : 1));
    L.beta = o;
    o += align16(8 * nrep);
    L.logu = o;
    o += align16(8 * nrep);
    L.E = o;
    o += align16(4 * nrep);
    L.s2r = o;
    o += align16(4 * nrep);
    L.r2s = o;
    o += align16(4 * nrep);
    L.att = o;
    o += align16(4 * nrep);
    L.succ = o;
    o += align16(4 * nrep);
    L.key = o;
    o += align16(4 * nrep);
    L.perm = o;
    o += align16(4 * nrep);
    L.tpos = o;
    o += align16(4 * nrep);
    L.level = o;
    o += align16(4 * nrep);
    const int nw = (n + 31) / 32;
    const int sweep_bytes = spin_bytes(nw);
    const int icm_bytes = icm ? icm_scratch_bytes(n) : 0;
    L.scratch_stride = align16(sweep_bytes > icm_bytes ? sweep_bytes : icm_bytes);
    L.scratch = o;
    o += W * L.scratch_stride;
    L.total = o;
    return L;
}

template <int Fmt, int W, int G>
__global__ __launch_bounds__(W* kWave) void k_fused(DevArgs a, int nsteps) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int tid = threadIdx.x;
    const int inst = a.order[blockIdx.x];
    const InstDesc d = a.desc[inst];
    const int n = d.n, nw = d.nw, nt = d.nt, nrep = d.nrep;
    const int adj16 = (int)d.adj_len16;
    const FusedLayout L = fused_layout(n, adj16, nrep, nt, d.fmax, W, G ? 0 : 1, d.nc > 0 ? 1 : 0);
    uint32_t* info = reinterpret_cast<uint32_t*>(smem + L.info);
    uint32_t* thr = reinterpret_cast<uint32_t*>(smem + L.thr);
    double* beta = reinterpret_cast<double*>(smem + L.beta);
    int32_t* E = reinterpret_cast<int32_t*>(smem + L.E);
    int32_t* s2r = reinterpret_cast<int32_t*>(smem + L.s2r);
    int32_t* r2s = reinterpret_cast<int32_t*>(smem + L.r2s);
    uint32_t* att = reinterpret_cast<uint32_t*>(smem + L.att);
    uint32_t* succ = reinterpret_cast<uint32_t*>(smem + L.succ);
    ExchScratch xs;
    xs.logu = reinterpret_cast<double*>(smem + L.logu);
    xs.key = reinterpret_cast<uint32_t*>(smem + L.key);
    xs.perm = reinterpret_cast<int32_t*>(smem + L.perm);
    xs.tpos = reinterpret_cast<int32_t*>(smem + L.tpos);
    xs.level = reinterpret_cast<int32_t*>(smem + L.level);
    uint8_t* scratch = smem + L.scratch + wave * L.scratch_stride;

    // ---- stage the instance ----
    if (!G) {
        stage_to_lds(info, a.rowinfo + d.pos_off, n, tid, W * kWave);
        stage_to_lds(reinterpret_cast<uint4*>(smem + L.adj), reinterpret_cast<const uint4*>(a.adj + d.adj_off16 * 16),
                     adj16, tid, W * kWave);
    }
    stage_to_lds(thr, a.thr + d.thr_off, nt * d.fmax, tid, W * kWave);
    const int64_t ro = d.rep_off;
    for (int s = tid; s < nrep; s += W * kWave) {
        beta[s] = a.beta[d.beta_off + s];
        E[s] = a.E[ro + s];
        s2r[s] = a.slot2rep[ro + s];
        r2s[s] = a.rep2slot[ro + s];
        att[s] = 0;
        succ[s] = 0;
    }
    __syncthreads();
    const uint32_t sent = sentinel_pos(d.nw);
    const Csr<Fmt> csr = G ? Csr<Fmt>{a.rowinfo + d.pos_off, a.adj + d.adj_off16 * 16, d.neg, d.kbit, 0, sent}
                           : Csr<Fmt>{info, smem + L.adj, d.neg, d.kbit, 0, sent};
    const int32_t* hf = a.hfield + d.pos_off;
    const int32_t* cls = a.cls + d.cls_off;
    const uint32_t c3s = ctr3(kTagSweep | a.tag_base, d.inst_id);
    const uint32_t c3i = ctr3(kTagIcm | a.tag_base, d.inst_id);
    const uint32_t step0 = a.inst_step[inst];
    unsigned long long nsw_total = 0;
    IcmScratch sc;
    sc.S1 = reinterpret_cast<uint32_t*>(scratch);
    sc.S2 = sc.S1 + (nw + 1);
    sc.D = sc.S2 + (nw + 1);
    sc.V = sc.D + (nw + 1);
    sc.tail = reinterpret_cast<int32_t*>(sc.V + (nw + 1));
    sc.Q = reinterpret_cast<uint16_t*>(sc.tail + 4);
    sc.red = nullptr;  // one wave per pair: no workgroup sums
    sc.Do = sc.Vo = nullptr;  // queue growth only

    const bool prof = a.phase != nullptr && blockIdx.x == 0 && tid == 0;
    unsigned long long t_mark = prof ? wall_clock64() : 0ull;
    auto phase_done = [&](int k) {
        if (prof) {
            const unsigned long long t = wall_clock64();
            a.phase[k] += t - t_mark;
            t_mark = t;
        }
    };
    for (int it = 0; it < nsteps; ++it) {
        const uint32_t step = step0 + (uint32_t)it;
        // ---- sweeps: replica r = wave, wave + W, ... ----
        for (int r = wave; r < nrep; r += W) {
            const int64_t gr = ro + r;
            uint32_t* words = a.spins + d.spin_off + (int64_t)r * nw;
            uint32_t* best_words = a.best_spins + d.spin_off + (int64_t)r * nw;
            const uint32_t* thr_row = thr + (r2s[r] % nt) * d.fmax;
            expand_bits(words, nw, scratch, lane);
            SweepState st;
            st.E = E[r];
            st.bestE = a.bestE[gr];
            st.best_sweep = a.best_sweep[gr];
            st.flips = 0;
            wave_sync();
            for (int sw = 0; sw < a.sweeps; ++sw) {
                const uint32_t gsweep = step * (uint32_t)a.sweeps + (uint32_t)sw;
                sweep_replica(csr, cls, d.ncls, hf, d.has_field != 0, nw, c3s, a.seed, r, gsweep, scratch, thr_row,
                              best_words, st, lane);
            }
            pack_bits(scratch, nw, words, lane);
            if (lane == 0) {
                E[r] = st.E;
                a.bestE[gr] = st.bestE;
                a.best_sweep[gr] = st.best_sweep;
                a.flips[gr] += st.flips;
            }
            wave_sync();
        }
        __syncthreads();
        phase_done(0);
        // ---- histogram + exchange (wave 0) ----
        if (wave == 0) {
            record_histogram(a, d, step, E, s2r, lane);
            if (!a.no_exchange) {
                nsw_total += exchange_wave(nt, nrep, a.pilot != 0, a.seed, a.tag_base, d.inst_id, step, beta, E, s2r,
                                           att, succ, xs, lane);
                wave_sync();
                for (int s = lane; s < nrep; s += kWave) r2s[s2r[s]] = s;
            }
        }
        __syncthreads();
        phase_done(1);
        // ---- cluster moves: slot t = nt-1-j, j = wave, wave + W, ... < nc ----
        if (d.nc > 0) {
            for (int j = wave; j < d.nc; j += W) {
                const int t = nt - 1 - j;
                const int r1 = s2r[t], r2 = s2r[t + nt];
                long long dE1 = 0, dE2 = 0;
                icm_team<Fmt, false>(csr, hf,
                                     d.has_field != 0, n, nw, t, step, a.seed, c3i,
                                     a.spins + d.spin_off + (int64_t)r1 * nw, a.spins + d.spin_off + (int64_t)r2 * nw,
                                     sc, dE1, dE2, lane, kWave);
                if (lane == 0) {
                    E[r1] += (int32_t)dE1;
                    E[r2] += (int32_t)dE2;
                }
                wave_sync();
            }
            __syncthreads();
        }
        phase_done(2);
    }
    // ---- write back ----
    for (int s = tid; s < nrep; s += W * kWave) {
        a.E[ro + s] = E[s];
        a.slot2rep[ro + s] = s2r[s];
        a.rep2slot[ro + s] = r2s[s];
        a.swap_att[ro + s] += att[s];
        a.swap_succ[ro + s] += succ[s];
    }
    if (wave == 0) {
        const unsigned long long tot = (unsigned long long)wave_sum_i64((long long)nsw_total);
        if (lane == 0) {
            a.swaps[inst] += tot;
            a.inst_step[inst] = step0 + (uint32_t)nsteps;
        }
    }
}

// Ends a split-mode PT step: every active instance's step index advances
// (a kernel rather than a host copy, so that graph replays advance it too).
// Per-slot observables at the end of a PT step, from step hist_start on
// (--log; the legacy writer of src/main.cc:490-530).  One wave per active
// instance; lane r takes replica r and adds to the slot it holds (the
// slots of an instance's replicas are distinct, so no two lanes share one).
__global__ __launch_bounds__(kWave) void k_observe(DevArgs a) {
    const int inst = a.order[blockIdx.x];
    if ((int64_t)a.inst_step[inst] < a.hist_start) return;
    const InstDesc d = a.desc[inst];
    const uint32_t last = (d.n & 31) ? (1u << (d.n & 31)) - 1u : 0xFFFFFFFFu;
    for (int r = threadIdx.x; r < d.nrep; r += kWave) {
        const uint32_t* w = a.spins + d.spin_off + (int64_t)r * d.nw;
        int up = 0;
        for (int k = 0; k + 1 < d.nw; ++k) up += __popc(w[k]);
        up += __popc(w[d.nw - 1] & last);
        const int64_t gs = d.rep_off + a.rep2slot[d.rep_off + r];
        const long long E = a.E[d.rep_off + r];
        long long* oi = a.obs_i + 4 * gs;
        double* od = a.obs_d + 2 * gs;
        oi[0] += 1;
        oi[1] += E;
        oi[2] += up;
        oi[3] += up * 2 > d.n ? up * 2 - d.n : d.n - up * 2;
        od[0] += (double)E * (double)E;
        od[1] += (double)up * (double)up;
    }
}

__global__ void k_step_advance(const int32_t* order, int n, uint32_t* inst_step) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) inst_step[order[t]] += 1u;
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

// Plain copy of n 16-byte units (calibrates the memory-traffic counters).
__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// LDS read bandwidth: every lane reads 16 bytes per iteration from a 32 KiB
// buffer (consecutive lanes, consecutive 16-byte slots) and folds them into
// one word so the loads are kept.  Bytes read = blocks * 256 * 16 * iters.
__global__ __launch_bounds__(256) void k_lds_read(uint32_t* out, int iters, uint32_t salt, int stride) {
    // 16-byte LDS reads at a runtime stride: the compiler cannot fold the
    // loop, and 4 independent accumulators keep 4 reads in flight per lane.
    __shared__ uint4 buf[2048];
    const int tid = threadIdx.x;
    for (int i = tid; i < 2048; i += 256) buf[i] = make_uint4(i * salt, i + salt, i ^ salt, i);
    __syncthreads();
    uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int idx = tid;
    for (int it = 0; it < iters; it += 4) {
        const uint4 v0 = buf[idx];
        const uint4 v1 = buf[(idx + stride) & 2047];
        const uint4 v2 = buf[(idx + 2 * stride) & 2047];
        const uint4 v3 = buf[(idx + 3 * stride) & 2047];
        a0 += v0.x + v1.y;
        a1 += v1.x + v2.y;
        a2 += v2.x + v3.y + v0.z + v2.w;
        a3 += v3.x + v0.y + v1.z + v3.w + v2.z + v3.z + v0.w + v1.w;
        idx = (idx + 4 * stride) & 2047;
    }
    const uint32_t acc = a0 ^ a1 ^ a2 ^ a3;
    if (acc == salt) out[blockIdx.x] = acc;
}

// explicit instantiations used by the host launcher
#define PTX_SPLIT_INST(F)                                          \
    template __global__ void k_init<F, 1>(DevArgs, int);          \
    template __global__ void k_init<F, 2>(DevArgs, int);          \
    template __global__ void k_init<F, 4>(DevArgs, int);          \
    template __global__ void k_init<F, 8>(DevArgs, int);          \
    template __global__ void k_sweep<F, 1>(DevArgs);              \
    template __global__ void k_sweep<F, 2>(DevArgs);              \
    template __global__ void k_sweep<F, 4>(DevArgs);              \
    template __global__ void k_sweep<F, 8>(DevArgs);              \
    template __global__ void k_icm<F, 1>(DevArgs, int, int);      \
    template __global__ void k_icm<F, 2>(DevArgs, int, int);      \
    template __global__ void k_icm<F, 4>(DevArgs, int, int);      \
    template __global__ void k_fused<F, 4, 0>(DevArgs, int);      \
    template __global__ void k_fused<F, 8, 0>(DevArgs, int);      \
    template __global__ void k_fused<F, 16, 0>(DevArgs, int);     \
    template __global__ void k_fused<F, 4, 1>(DevArgs, int);      \
    template __global__ void k_fused<F, 8, 1>(DevArgs, int);      \
    template __global__ void k_fused<F, 16, 1>(DevArgs, int);
PTX_SPLIT_INST(kFmtWide)
PTX_SPLIT_INST(kFmtUnit)
template __global__ void k_sweep_rs<1>(DevArgs);
template __global__ void k_sweep_rs<2>(DevArgs);
template __global__ void k_sweep_rs<4>(DevArgs);
template __global__ void k_sweep_rs<8>(DevArgs);

}  // namespace ptx
