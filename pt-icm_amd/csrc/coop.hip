// coop.hip -- persistent kernel for one instance or a small batch: one
// workgroup per replica of every instance in the launch, all co-resident
// (cooperative launch), several PT steps per launch.
//
// Per PT step each workgroup sweeps its replica with all of its waves, then
// publishes its spin words and tagged energy to a double-buffered area, then
// waits until every replica of its instance has published the step.  After
// that every workgroup computes the whole (NN)_a exchange itself from the
// published energies (the computation is deterministic, so every copy of
// the permutation is identical), and the two workgroups of each ICM slot
// pair both grow the same cluster from the published states and each keeps
// its own replica's half.  No second barrier is needed: a buffer is
// rewritten only two steps later, after every replica of the instance has
// published the step in between.
#include "kernels.hpp"

namespace ptx {

struct CoopArgs {
    uint32_t* pub_spins;        // [2][spin words of all instances], at each instance's spin offset
    unsigned long long* pub_E;  // [2][replicas of all instances]: (step + 1) << 32 | energy, at each
                                // instance's replica offset; a slot showing this step is the arrival
    const int2* items;          // (instance, replica) of each workgroup
    uint32_t* err;              // set to 1 when a wait gives up (read by the other workgroups)
    uint32_t* fault;            // host-mapped fault words: [0] set with err, for the host to read without a sync
    uint64_t wait_ticks;        // bound of one step wait in wall-clock ticks; 0: give up at the first
                                // poll that finds a replica missing (forces the error path in tests)
    int64_t pub_words;          // spin words of all instances (the stride of pub_spins' halves)
    int32_t pub_reps;           // replicas of all instances (the stride of pub_E's halves)
    int32_t pre_rnd;            // 1: the LDS layout holds the next sweep's random words
    int32_t lat_stage;          // 1: the lattice tables are copied into LDS
    int32_t pre_f0;             // 1: the LDS layout holds the next sweep's class-0 local fields (drawn
                                // during the step wait; kernels without cluster moves)
};

// Metropolis sweep of one replica by a whole workgroup (W waves).  Lane gl of
// nthr takes SPL positions per step of the class loop: SPL = 1 -> position
// cs + gl + k*nthr; SPL = 4 -> the aligned groups of 4 used by the wave-level
// sweep.  Every position draws word p&3 of the Philox block p>>2, so the
// result is the same for any SPL and workgroup size.  Each wave's class sums
// go to its slot in one of two alternating LDS buffers (one barrier per class).
constexpr int kMaxCoopWaves = 16;  // waves per cooperative workgroup (4, 8, or 16 without cluster moves)

// The random words of a whole sweep of replica r: pre[p] = word p&3 of the
// Philox block (p>>2, r, gsweep) -- the words the sweep would draw for p.
// Threads t0, t0 + stride, ... each take whole blocks.
__device__ __forceinline__ void fill_sweep_randoms(uint32_t* pre, int n, uint32_t c3, uint64_t seed, int r,
                                                   uint32_t gsweep, int t0, int stride) {
    for (int g = t0; 4 * g < n; g += stride) {
        const Philox4 x = philox4x32_10((uint32_t)g, (uint32_t)r, gsweep, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
        *reinterpret_cast<uint4*>(pre + 4 * g) = make_uint4(x.v[0], x.v[1], x.v[2], x.v[3]);
    }
}

// Local fields (with the field h) of the positions [cs, ce) of one class for
// the next sweep, by threads t0, t0 + stride, ...
template <int Fmt>
__device__ __forceinline__ void class_fields(const Csr<Fmt>& csr, int cs, int ce, const int32_t* __restrict__ hf,
                                             bool has_field, const uint8_t* sp, int32_t* f0, int t0, int stride) {
    for (int p = cs + t0; p < ce; p += stride) f0[p - cs] = csr.field(p, sp) + (has_field ? hf[p] : 0);
}

template <int Fmt, int SPL>
__device__ __forceinline__ void sweep_replica_wg(const Csr<Fmt>& csr, const int32_t* __restrict__ cls, int ncls,
                                                 const int32_t* __restrict__ hf, bool has_field, int nw, uint32_t c3,
                                                 uint64_t seed, int r, uint32_t gsweep, uint8_t* sp,
                                                 const uint32_t* thr_row, uint32_t* best_words, SweepState& st,
                                                 int gl, int nthr, int* acc, uint32_t& cc, const uint32_t* pre,
                                                 const int32_t* f0) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const int lane = gl % kWave;
    const int nwv = nthr / kWave, wv = gl / kWave;
    for (int c = 0; c < ncls; ++c, ++cc) {
        // cc counts classes across sweeps and steps: each wave stores its
        // (dE, flips) sums in slot wv of buffer cc % 2; a buffer is read right
        // after its class's barrier and rewritten only after the next one.
        const int cs = cls[c], ce = cls[c + 1];
        int dsum = 0, nacc = 0;
        if (c == 0 && f0) {
            // local fields drawn during the previous step's wait: decisions only
            for (int p = cs + gl; p < ce; p += nthr) {
                const uint32_t b = sp[p];
                const int f = f0[p - cs];
                const int dE = b ? -2 * f : 2 * f;
                bool accept = dE <= 0;
                if (!accept) {
                    const uint32_t u =
                        pre ? pre[p]
                            : pick(philox4x32_10((uint32_t)p >> 2, (uint32_t)r, gsweep, c3, k0, k1), (uint32_t)p);
                    accept = u < thr_row[(dE >> 1) - 1];
                }
                if (accept) {
                    sp[p] = b ^ 1u;
                    dsum += dE;
                }
                nacc += (int)__popcll(__ballot(accept));
            }
        } else if (SPL == 1) {
            // T lanes per position when the class has at most nthr / T of
            // them (T = 2, 4): each lane gathers every T-th chunk of the row
            // and a quad permutation sums the parts, so short classes use
            // every wave and each lane's gather chain is shorter.
            auto update = [&](auto tc) {
                constexpr int T = decltype(tc)::value;
                const int sub = gl & (T - 1);
                for (int p = cs + gl / T; p < ce; p += nthr / T) {
                    // the position's own byte and random word are read before
                    // the gathers, off the decision's chain of LDS reads
                    const uint32_t b = sp[p];
                    const uint32_t u0 = pre ? pre[p] : 0u;
                    int f;
                    if (T == 1) {
                        f = csr.field(p, sp);
                    } else {
                        int d;
                        const int part = group_sum_i32<T>(csr.template field_part<T>(p, sp, sub, d));
                        f = Fmt == kFmtUnit ? 2 * part - d : part;
                    }
                    f += has_field ? hf[p] : 0;
                    bool took = false;
                    if (sub == 0) {
                        const int dE = b ? -2 * f : 2 * f;
                        bool accept = dE <= 0;
                        if (!accept) {
                            const uint32_t u =
                                pre ? u0
                                    : pick(philox4x32_10((uint32_t)p >> 2, (uint32_t)r, gsweep, c3, k0, k1),
                                           (uint32_t)p);
                            accept = u < thr_row[(dE >> 1) - 1];
                        }
                        if (accept) {
                            sp[p] = b ^ 1u;
                            dsum += dE;
                            took = true;
                        }
                    }
                    nacc += (int)__popcll(__ballot(took));  // wave-uniform count, no reduction needed
                }
            };
            const int csz = ce - cs;
            if (4 * csz <= nthr)
                update(std::integral_constant<int, 4>{});
            else if (2 * csz <= nthr)
                update(std::integral_constant<int, 2>{});
            else
                update(std::integral_constant<int, 1>{});
        } else {
            for (int g = (cs >> 2) + gl; 4 * g < ce; g += nthr)
                group_update(csr, g, cs, ce, hf, has_field, c3, k0, k1, r, gsweep, sp, thr_row, dsum, nacc, pre);
        }
        dsum = wave_sum_i32(dsum);
        if (SPL != 1 && !(c == 0 && f0)) nacc = wave_sum_i32(nacc);
        int* a = acc + 2 * kMaxCoopWaves * (int)(cc & 1u);
        if (lane == 0) {
            a[2 * wv] = dsum;
            a[2 * wv + 1] = nacc;
        }
        __syncthreads();
        int ds = 0, na = 0;
        for (int q = 0; q < nwv; ++q) {
            ds += a[2 * q];
            na += a[2 * q + 1];
        }
        st.E += ds;
        st.flips += (unsigned long long)na;
        if (na > 0 && st.E < st.bestE) {  // uniform: every thread holds the same sums
            st.bestE = st.E;
            st.best_sweep = (int32_t)gsweep;
            pack_bits(sp, nw, best_words, gl, nthr);
            __syncthreads();  // the snapshot is complete before the next class flips spins
        }
    }
}

// Per-step barrier among the workgroups of one instance (other instances
// need not wait: no data crosses instances), in two halves so that waves
// other than the first can do independent work meanwhile.  Arrival: after
// the workgroup has published its spin words, thread 0 stores the tagged
// energy (step + 1) << 32 | E with release semantics at agent scope.  Wait:
// the first wave polls the instance's slots with relaxed loads (lane s reads
// slot s) until every one shows this step, takes the energies from the
// values it polled, and then one acquire fence makes the published spins
// visible.  Without cluster moves no workgroup reads another's spins, and
// the 64-bit tagged word carries the energy itself: REL = ACQ = false then
// store the slot relaxed, with no barrier before it and no fence after the
// wait.  The slots are double-buffered by step parity: a slot is
// rewritten two steps later, after every replica of the instance has
// published the step in between, which it does only after reading this one.
// coop_wait (wave 0 alone) returns false when the wait exceeded its bound of
// wait_ticks wall-clock ticks (100 MHz), or another workgroup gave up: the
// error flag is then set, the host-mapped fault word too, and every
// workgroup leaves the step loop, so the launch always ends.
template <bool REL>
__device__ __forceinline__ void coop_arrive(unsigned long long* slot, uint32_t step, int32_t E) {
    if (REL) __syncthreads();  // the spin words are published
    if (threadIdx.x == 0)
        __hip_atomic_store(slot, ((unsigned long long)(step + 1u) << 32) | (uint32_t)E,
                           REL ? __ATOMIC_RELEASE : __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane l polls slot slot0 in the first round (slot0 = l, or a permutation
// of 0..nrep-1 over the lanes below nrep) and returns that slot's energy in
// e0; E[s] receives every slot's energy.
template <bool ACQ>
__device__ __forceinline__ bool coop_wait(const unsigned long long* slots, int nrep, uint32_t step, const CoopArgs& c,
                                          int32_t* E, int slot0, int32_t& e0) {
    const int tid = threadIdx.x;
    int ok = 1;
    e0 = 0;
    {
        const unsigned long long want = (unsigned long long)(step + 1u);
        unsigned long long spins = 0;
        const uint64_t t0 = wall_clock64();
        for (int s0 = 0; s0 < nrep; s0 += kWave) {
            const int s = s0 == 0 ? slot0 : s0 + tid;
            unsigned long long v = 0;
            while (true) {
                bool ready = true;
                if (s < nrep) {
                    v = __hip_atomic_load(slots + s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ready = (v >> 32) == want;
                }
                if (__ballot(!ready) == 0ull) break;
                __builtin_amdgcn_s_sleep(1);
                // the error flag and the time bound every 32 polls (a poll is one memory round trip)
                if (((++spins & 31u) == 0 || c.wait_ticks == 0) &&
                    (wall_clock64() - t0 > c.wait_ticks ||
                     __hip_atomic_load(c.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    if (tid == 0) {
                        atomicOr(c.err, 1u);
                        __hip_atomic_store(c.fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    }
                    ok = 0;
                    break;
                }
            }
            if (!ok) break;
            if (s < nrep) E[s] = (int32_t)(uint32_t)v;
            if (s0 == 0) e0 = (int32_t)(uint32_t)v;
        }
        if (ACQ) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    return ok != 0;
}

struct CoopLayout {
    int info, adj, cls, thr, beta, E, s2r, r2s, att, succ, logu, key, perm, tpos, level, pre, sp, icm, lat, red,
        lat_tab, acc, flag, lua, kbuf, rnd, f0, total;
};

// pre_rnd = 1 adds the random words of one sweep (4 bytes per position);
// lat_words > 0 two lattice-order bit sets for cluster growth, and with
// lat_stage = 1 also a copy of the lattice tables (wpos, obit); csr_global
// = 1 leaves the adjacency in global memory (read through L2), so that more
// workgroups fit a CU; f0_len > 0 adds the next sweep's local fields of
// colour class 0 (f0_len positions, 4 bytes each).
__host__ __device__ inline CoopLayout coop_layout(int n, int adj_len16, int ncls, int nrep, int nt, int fmax,
                                                  int pre_rnd, int lat_words, int lat_stage, int csr_global,
                                                  int f0_len = 0) {
    CoopLayout L;
    int o = 0;
    auto take = [&](int bytes) {
        const int at = o;
        o += (bytes + 15) & ~15;
        return at;
    };
    const int nw = (n + 31) / 32;
    L.sp = take(spin_bytes(nw));  // at LDS address 0: a gather's address is the record itself
    L.info = take(csr_global ? 0 : 4 * n);
    L.adj = take(csr_global ? 0 : 16 * (adj_len16 > 0 ? adj_len16 : 1));
    L.cls = take(4 * (ncls + 1));
    L.thr = take(4 * nt * (fmax > 0 ? fmax : 1));
    L.beta = take(8 * nrep);
    L.logu = take(8 * nrep);
    L.E = take(4 * nrep);
    L.s2r = take(4 * nrep);
    L.r2s = take(4 * nrep);
    L.att = take(4 * nrep);
    L.succ = take(4 * nrep);
    L.key = take(4 * nrep);
    L.perm = take(4 * nrep);
    L.tpos = take(4 * nrep);
    L.level = take(4 * nrep);
    L.pre = take(2 * kWave * (int)sizeof(ExchPre));  // prepared exchange of the next step, two buffers
    L.icm = take(icm_scratch_bytes(n));
    L.lat = take(8 * lat_words);
    L.red = take(8 * kIcmRedSlots);
    L.lat_tab = take(lat_stage && lat_words > 0 ? 4 * (16 * lat_words + n) : 0);
    L.acc = take(4 * 2 * 2 * kMaxCoopWaves);
    L.flag = take(8);  // [0] an energy shared by the waves, [1] the step's wait succeeded
    L.lua = take(2 * kWave * 8);  // log-uniforms of the exchange attempts, by step parity
    L.kbuf = take(2 * kWave * 4);
    L.rnd = take(pre_rnd ? 16 * ((n + 3) / 4) : 0);
    L.f0 = take(4 * f0_len);
    L.total = o;
    return L;
}

// IC = 0: no instance of the batch has cluster-move slots, and the kernel
// carries no cluster-growth code (whose registers would otherwise be
// allocated around the sweep as well).
template <int Fmt, int W, int SPL, int G, int IC>
__global__ __launch_bounds__(W* kWave) void k_coop(DevArgs a, CoopArgs c, int nsteps) {
    extern __shared__ __align__(16) uint8_t smem[];
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    const int nthr = W * kWave;
    const int2 item = c.items[blockIdx.x];
    const int inst = item.x;
    const InstDesc d = a.desc[inst];
    const int n = d.n, nw = d.nw, nt = d.nt, nrep = d.nrep;
    const int r = item.y;  // this workgroup's replica of its instance
    const CoopLayout L = coop_layout(n, (int)d.adj_len16, d.ncls, nrep, nt, d.fmax, c.pre_rnd, d.lat_words,
                                     c.lat_stage, G, c.pre_f0 ? d.cls0 : 0);
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
    uint8_t* sp = smem + L.sp;
    int* acc = reinterpret_cast<int*>(smem + L.acc);
    int* flag = reinterpret_cast<int*>(smem + L.flag);
    IcmScratch sc;
    sc.S1 = reinterpret_cast<uint32_t*>(smem + L.icm);
    sc.S2 = sc.S1 + (nw + 1);
    sc.D = sc.S2 + (nw + 1);
    sc.V = sc.D + (nw + 1);
    sc.tail = reinterpret_cast<int32_t*>(sc.V + (nw + 1));
    sc.Q = reinterpret_cast<uint16_t*>(sc.tail + 4);
    sc.red = reinterpret_cast<long long*>(smem + L.red);
    sc.Do = reinterpret_cast<uint32_t*>(smem + L.lat);
    sc.Vo = sc.Do + d.lat_words;

    // a wait of an earlier launch gave up: leave at once (one load, shared
    // through LDS so that every wave takes the same branch)
    if (tid == 0) *flag = (int)__hip_atomic_load(c.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (*flag) return;

    // ---- stage the instance and this replica ----
    if (!G) {  // G = 1: the adjacency is read through L2
        stage_to_lds(info, a.rowinfo + d.pos_off, n, tid, nthr);
        stage_to_lds(reinterpret_cast<uint4*>(smem + L.adj), reinterpret_cast<const uint4*>(a.adj + d.adj_off16 * 16),
                     (int)d.adj_len16, tid, nthr);
    }
    stage_to_lds(thr, a.thr + d.thr_off, nt * d.fmax, tid, nthr);
    int32_t* cls = reinterpret_cast<int32_t*>(smem + L.cls);  // class bounds, read at every class
    for (int m = tid; m <= d.ncls; m += nthr) cls[m] = a.cls[d.cls_off + m];
    LatView lv = lat_view(a, d);
    if (c.lat_stage && d.lat_dim > 0) {  // the tables of this lattice shape, read by every cluster move
        uint32_t* tab = reinterpret_cast<uint32_t*>(smem + L.lat_tab);
        const uint32_t* src = a.lat + d.lat_off;
        for (int m = tid; m < 16 * d.lat_words + n; m += nthr) tab[m] = src[m];
        lv.wpos = reinterpret_cast<const uint16_t*>(tab);
        lv.obit = tab + 16 * d.lat_words;
    }
    const int64_t ro = d.rep_off;
    for (int s = tid; s < nrep; s += nthr) {
        beta[s] = a.beta[d.beta_off + s];
        s2r[s] = a.slot2rep[ro + s];
        r2s[s] = a.rep2slot[ro + s];
        att[s] = 0;
        succ[s] = 0;
    }
    if (tid < 4 * kMaxCoopWaves) acc[tid] = 0;
    expand_bits(a.spins + d.spin_off + (int64_t)r * nw, nw, sp, tid, nthr);
    // register-resident exchange (up to 64 replicas): the first step's
    // preparation here, later ones overlapped with the previous step's wait
    // and exchange (waves 1 and 2 prepare one half each and store their
    // fields of ExchPre; with eight or more waves wave 3 draws the attempts'
    // log-uniforms meanwhile)
    const bool prepared = nrep <= kWave && !a.no_exchange && nt > 1;
    const bool lu_apart = prepared && W >= 8;
    ExchPre* pre = reinterpret_cast<ExchPre*>(smem + L.pre);
    double* lua = reinterpret_cast<double*>(smem + L.lua);
    uint32_t* kbuf = reinterpret_cast<uint32_t*>(smem + L.kbuf);  // sort keys, one block of 64 per preparing wave
    auto prepare_half = [&](uint32_t s, int h) {
        const ExchPre x = exch_prepare(nt, nrep, a.pilot != 0, a.seed, a.tag_base, d.inst_id, s, lane,
                                       kbuf + h * kWave, h, !lu_apart);
        ExchPre& slot = pre[(s & 1u) * kWave + lane];
        // constant indices keep x in registers
        if (h == 0) {
            slot.tpos[0] = x.tpos[0];
            slot.level[0] = x.level[0];
            slot.lu[0] = x.lu[0];
            slot.maxlv[0] = x.maxlv[0];
        } else {
            slot.tpos[1] = x.tpos[1];
            slot.level[1] = x.level[1];
            slot.lu[1] = x.lu[1];
            slot.maxlv[1] = x.maxlv[1];
        }
    };
    if (prepared && (wave == 1 || wave == 2)) prepare_half(a.inst_step[inst], wave - 1);
    if (lu_apart && wave == 3)
        exch_attempt_logu(nt, a.pilot != 0, a.seed, a.tag_base, d.inst_id, a.inst_step[inst], lane,
                          lua + (a.inst_step[inst] & 1u) * kWave);
    // random words of the first sweep of this launch (later ones are drawn
    // while the step's wait is pending)
    uint32_t* pre_rnd = c.pre_rnd ? reinterpret_cast<uint32_t*>(smem + L.rnd) : nullptr;
    int32_t* f0buf = !IC && c.pre_f0 ? reinterpret_cast<int32_t*>(smem + L.f0) : nullptr;
    const uint32_t c3s0 = ctr3(kTagSweep | a.tag_base, d.inst_id);
    if (pre_rnd) fill_sweep_randoms(pre_rnd, n, c3s0, a.seed, r, a.inst_step[inst] * (uint32_t)a.sweeps, tid, nthr);
    __syncthreads();
    const uint32_t sent = sentinel_pos(d.nw);
    const Csr<Fmt> csr = G ? Csr<Fmt>{a.rowinfo + d.pos_off, a.adj + d.adj_off16 * 16, d.neg, d.kbit, 0, sent}
                           : Csr<Fmt>{info, smem + L.adj, d.neg, d.kbit, 0, sent};
    const int32_t* hf = a.hfield + d.pos_off;
    const uint32_t c3s = ctr3(kTagSweep | a.tag_base, d.inst_id);
    const uint32_t c3i = ctr3(kTagIcm | a.tag_base, d.inst_id);
    const uint32_t step0 = a.inst_step[inst];
    // this instance's halves of the published states and energies
    uint32_t* pub_sp = c.pub_spins + d.spin_off;
    unsigned long long* pub_e = c.pub_E + ro;
    const int64_t gr = ro + r;
    uint32_t* best_words = a.best_spins + d.spin_off + (int64_t)r * nw;
    SweepState st;
    st.E = a.E[gr];
    st.bestE = a.bestE[gr];
    st.best_sweep = a.best_sweep[gr];
    st.flips = 0;
    unsigned long long nsw_total = 0;
    uint32_t cc = 0;  // running class counter of the accumulator rotation (acc[] is zero at start)
    const bool prof = a.phase != nullptr && blockIdx.x == 0 && tid == 0;
    unsigned long long t_mark = prof ? wall_clock64() : 0ull;
    auto phase_done = [&](int k) {
        if (prof) {
            const unsigned long long t = wall_clock64();
            a.phase[k] += t - t_mark;
            t_mark = t;
        }
    };
    int it = 0;
    for (; it < nsteps; ++it) {
        const uint32_t step = step0 + (uint32_t)it;
        const int par = (int)(step & 1u);
        // ---- sweep ----
        const uint32_t* thr_row = thr + (r2s[r] % nt) * d.fmax;
        for (int sw = 0; sw < a.sweeps; ++sw) {
            const uint32_t gsweep = step * (uint32_t)a.sweeps + (uint32_t)sw;
            sweep_replica_wg<Fmt, SPL>(csr, cls, d.ncls, hf, d.has_field != 0, nw, c3s, a.seed, r, gsweep, sp,
                                       thr_row, best_words, st, tid, nthr, acc, cc, sw == 0 ? pre_rnd : nullptr,
                                       sw == 0 && it > 0 ? f0buf : nullptr);
        }
        phase_done(0);
        // ---- publish and wait ----
        // the published states are read only by cluster moves; without them
        // the tagged energy slot is all a replica publishes (a relaxed
        // agent-scope store: no release, no acquire after the wait)
        if (IC) pack_bits(sp, nw, pub_sp + par * c.pub_words + (int64_t)r * nw, tid, nthr);
        coop_arrive<IC != 0>(pub_e + (int64_t)par * c.pub_reps + r, step, st.E);
        // Wave 0 waits for the step's energies, then records the histogram
        // (workgroup 0) and computes the exchange (every workgroup).  The
        // other waves meanwhile prepare the next step: waves 1 and 2 its
        // exchange, wave 3 (eight or more waves) its attempts' log-uniforms,
        // and the waves after them (waves 1.. without a preparation) its
        // first-sweep random words (this step's sweep has finished reading
        // them).  One barrier closes the step; flag[1] tells every wave
        // whether the wait succeeded.
        if (prepared && (wave == 1 || wave == 2) && it + 1 < nsteps) prepare_half(step + 1u, wave - 1);
        if (lu_apart && wave == 3 && it + 1 < nsteps)
            exch_attempt_logu(nt, a.pilot != 0, a.seed, a.tag_base, d.inst_id, step + 1u, lane,
                              lua + ((step + 1u) & 1u) * kWave);
        const int w0 = lu_apart ? 4 : 1;
        if (pre_rnd && wave >= w0 && it + 1 < nsteps)
            fill_sweep_randoms(pre_rnd, n, c3s, a.seed, r, (step + 1u) * (uint32_t)a.sweeps, tid - w0 * kWave,
                               nthr - w0 * kWave);
        // without cluster moves the state stays as this sweep left it until
        // the next one: its class-0 local fields are drawn now
        if (f0buf && wave >= w0 && it + 1 < nsteps)
            class_fields(csr, cls[0], cls[1], hf, d.has_field != 0, sp, f0buf, tid - w0 * kWave, nthr - w0 * kWave);
        if (wave == 0) {
            // the prepared exchange's inputs that do not depend on this
            // step's energies are read before the wait; each lane polls the
            // slot of the replica its temperature slot holds, so that its
            // energy arrives in the register the exchange starts from
            ExchPre x;
            int rep0 = lane;
            double bk = 0.0;
            if (prepared) {
                x = pre[(step & 1u) * kWave + lane];
                if (lu_apart) {
                    const double* lu = lua + (step & 1u) * kWave;
                    x.lu[0] = x.tpos[0] >= 0 ? lu[x.tpos[0]] : 0.0;
                    x.lu[1] = x.tpos[1] >= 0 ? lu[x.tpos[1]] : 0.0;
                }
                if (lane < nrep) {
                    rep0 = s2r[lane];
                    bk = beta[lane];
                }
            }
            int32_t e0 = 0;
            const bool ok = coop_wait<IC != 0>(pub_e + (int64_t)par * c.pub_reps, nrep, step, c, E, rep0, e0);
            wave_sync();  // E as the lanes stored it, for the lanes that read it
            if (lane == 0) flag[1] = ok ? 1 : 0;
            phase_done(3);
            if (ok) {
                if (r == 0) record_histogram(a, d, step, E, s2r, lane);
                if (!a.no_exchange) {
                    if (prepared) {
                        // one lane per slot: r2s written from the lane's own registers
                        nsw_total += exch_apply(nt, nrep, a.pilot != 0, x, lane < nrep ? rep0 : 0,
                                                lane < nrep ? e0 : 0, bk, s2r, att, succ, lane, r2s);
                    } else {
                        nsw_total += exchange_wave(nt, nrep, a.pilot != 0, a.seed, a.tag_base, d.inst_id, step,
                                                   beta, E, s2r, att, succ, xs, lane);
                        wave_sync();
                        for (int s = lane; s < nrep; s += kWave) r2s[s2r[s]] = s;
                    }
                }
            }
        }
        __syncthreads();
        if (!flag[1]) break;
        phase_done(1);
        // ---- cluster move of this replica's slot pair (both partners compute it) ----
        if (IC && d.nc > 0) {
            const int t = r2s[r] % nt;
            if (t >= nt - d.nc) {
                const int r1 = s2r[t], r2 = s2r[t + nt];
                if (d.lat_dim == 0 && d.deg_mean >= kIcmBlockDeg) {
                    // wide levels: the whole workgroup grows the cluster and sums
                    // its boundary; every thread gets both sums
                    long long dE1 = 0, dE2 = 0;
                    icm_team<Fmt, true>(csr, hf, d.has_field != 0, n, nw, t, step, a.seed, c3i,
                                        pub_sp + par * c.pub_words + (int64_t)r1 * nw,
                                        pub_sp + par * c.pub_words + (int64_t)r2 * nw, sc, dE1, dE2, tid, nthr,
                                        false, LatView{}, true, E[r1], E[r2]);
                    st.E += (int32_t)(r == r1 ? dE1 : dE2);
                    expand_bits(r == r1 ? sc.S1 : sc.S2, nw, sp, tid, nthr);
                    __syncthreads();
                } else {
                    // lattices and sparse graphs: one wave grows the cluster (its
                    // frontiers are mostly narrower than a wave, so more waves
                    // would only add barriers)
                    if (wave == 0) {
                        long long dE1 = 0, dE2 = 0;
                        icm_team<Fmt, false>(csr, hf,
                                             d.has_field != 0, n, nw, t, step, a.seed, c3i,
                                             pub_sp + par * c.pub_words + (int64_t)r1 * nw,
                                             pub_sp + par * c.pub_words + (int64_t)r2 * nw, sc, dE1, dE2, lane, kWave,
                                             false, lv, true, E[r1], E[r2]);
                        if (lane == 0) flag[0] = st.E + (int32_t)(r == r1 ? dE1 : dE2);
                    }
                    __syncthreads();
                    st.E = flag[0];  // the new energy, shared with every wave
                    expand_bits(r == r1 ? sc.S1 : sc.S2, nw, sp, tid, nthr);
                    __syncthreads();
                }
            }
        }
        phase_done(2);
    }
    // ---- write back ----
    pack_bits(sp, nw, a.spins + d.spin_off + (int64_t)r * nw, tid, nthr);
    if (tid == 0) {
        a.E[gr] = st.E;
        a.bestE[gr] = st.bestE;
        a.best_sweep[gr] = st.best_sweep;
        a.flips[gr] += st.flips;
    }
    if (r == 0) {
        for (int s = tid; s < nrep; s += nthr) {
            a.slot2rep[ro + s] = s2r[s];
            a.rep2slot[ro + s] = r2s[s];
            a.swap_att[ro + s] += att[s];
            a.swap_succ[ro + s] += succ[s];
        }
        if (wave == 0) {
            const unsigned long long tot = (unsigned long long)wave_sum_i64((long long)nsw_total);
            if (lane == 0) {
                a.swaps[inst] += tot;
                a.inst_step[inst] = step0 + (uint32_t)it;
            }
        }
    }
}

#define PTX_COOP_INST(F, IC)                                          \
    template __global__ void k_coop<F, 4, 1, 0, IC>(DevArgs, CoopArgs, int); \
    template __global__ void k_coop<F, 8, 1, 0, IC>(DevArgs, CoopArgs, int); \
    template __global__ void k_coop<F, 4, 4, 0, IC>(DevArgs, CoopArgs, int); \
    template __global__ void k_coop<F, 8, 4, 0, IC>(DevArgs, CoopArgs, int); \
    template __global__ void k_coop<F, 4, 1, 1, IC>(DevArgs, CoopArgs, int);
PTX_COOP_INST(kFmtWide, 0)
PTX_COOP_INST(kFmtWide, 1)
PTX_COOP_INST(kFmtUnit, 0)
PTX_COOP_INST(kFmtUnit, 1)
// 16 waves (4 per SIMD: more LDS reads in flight per class), one position
// per lane, adjacency in LDS, no cluster moves
template __global__ void k_coop<kFmtWide, 16, 1, 0, 0>(DevArgs, CoopArgs, int);
template __global__ void k_coop<kFmtUnit, 16, 1, 0, 0>(DevArgs, CoopArgs, int);

}  // namespace ptx
