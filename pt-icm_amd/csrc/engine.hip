// engine.hip -- host orchestration of the gfx950 kernels (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <atomic>
#include <memory>
#include <thread>
#include <unordered_map>
#include <cstring>
#include <string>
#include <type_traits>

#include "kernels.hip"
#include "coop.hip"
#include "ladder.hpp"

namespace ptx {

Status hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return Status::ok();
    const int code = (e == hipErrorOutOfMemory) ? PTX_ERR_NOMEM : PTX_ERR_DEVICE;
    return Status::err(code, std::string(what) + ": " + hipGetErrorString(e));
}


template <class T>
Status DevBuf<T>::alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    PTX_TRY(hip_check(hipMalloc(&p, count * sizeof(T)), "hipMalloc"));
    n = count;
    // The fill runs on the null stream, which does not order against the
    // non-blocking streams engines and networks use: wait for it here so a
    // later copy or kernel on another stream cannot be overwritten by it.
    PTX_TRY(hip_check(hipMemset(p, 0, count * sizeof(T)), "hipMemset"));
    return hip_check(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
}
template <class T>
Status DevBuf<T>::upload(const std::vector<T>& v, hipStream_t s) {
    PTX_TRY(alloc(v.size()));
    if (v.empty()) return Status::ok();
    return hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}
template <class T>
Status DevBuf<T>::upload(const T* v, size_t count, size_t offset, hipStream_t s) {
    if (offset + count > n) return Status::err(PTX_ERR_INVALID, "upload out of range");
    return hip_check(hipMemcpyAsync(p + offset, v, count * sizeof(T), hipMemcpyHostToDevice, s), "upload");
}
template <class T>
Status DevBuf<T>::download(T* v, size_t count, size_t offset, hipStream_t s) const {
    if (offset + count > n) return Status::err(PTX_ERR_INVALID, "download out of range");
    PTX_TRY(hip_check(hipMemcpyAsync(v, p + offset, count * sizeof(T), hipMemcpyDeviceToHost, s), "download"));
    return hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

Engine::~Engine() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& t : pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    if (stream) (void)hipStreamDestroy(stream);
    if (h_fault) (void)hipHostFree(h_fault);
}

Status Engine::check_fault() const {
    if (!h_fault) return Status::ok();
    const uint32_t wait = __atomic_load_n(h_fault, __ATOMIC_ACQUIRE);
    const uint32_t hist = __atomic_load_n(h_fault + 1, __ATOMIC_ACQUIRE);
    if (wait)
        return Status::err(PTX_ERR_DEVICE, "a cooperative step wait gave up (a replica's workgroup did not arrive "
                                           "within the bound); the engine's state is no longer a valid trajectory");
    if (hist)
        return Status::err(PTX_ERR_DEVICE, std::to_string(hist) + " histogram samples fell outside their table");
    return Status::ok();
}

namespace {
int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Order of the records within each adjacency row (unit format), chosen for
// the cooperative sweep's neighbour gathers.  In a colour class of csz
// positions, T = 4, 2 or 1 lanes share a position (the first T with
// T * csz <= nthr) and lane sub of them reads the row's chunks sub, sub + T,
// ...; gather h of chunk ch is one LDS byte read of the whole wave, at the
// neighbour's position.  Two lanes of a wave that read different words of
// one LDS bank in the same gather serialise, and random neighbours make that
// common (about 3 words per bank and gather on G022).  Rows are walked in
// position order; each record takes, among its row's free slots, the one
// whose gather has the fewest other words in the record's bank so far in
// this wave.  The records stay in the row's first deg slots (code that walks
// a row's neighbours reads those; the padding after them reads the zero
// byte), and the local field is a sum over the row, so the bits do not change.
struct BankOrder {
    static constexpr int kBanks = 32;
    static constexpr int kKeys = 8 * (kMaxDegree / 8 + 1);  // gather (chunk / T, h) of a wave
    int nthr;
    std::vector<std::vector<uint32_t>> use;  // key * kBanks + bank -> words read so far by the wave
    std::vector<int> touched;
    int64_t wave_id = -1;
    explicit BankOrder(int threads) : nthr(threads), use((size_t)kKeys * kBanks) {}
    void order(int64_t cls_id, int j, int csz, uint16_t* rec, int deg, uint32_t pos_mask) {
        const int T = 4 * csz <= nthr ? 4 : (2 * csz <= nthr ? 2 : 1);
        const int per = nthr / T;
        const int64_t wid = (cls_id << 32) | ((int64_t)(j / per) << 8) | (int64_t)((j % per) * T / 64);
        if (wid != wave_id) {
            for (int t : touched) use[(size_t)t].clear();
            touched.clear();
            wave_id = wid;
        }
        const int S = deg;
        std::vector<uint16_t> in(rec, rec + S), out(S);
        std::vector<char> taken(S, 0);
        for (uint16_t r : in) {
            const uint32_t word = (r & pos_mask) >> 2, bank = word % kBanks;
            int best = -1, best_cost = 1 << 30;
            for (int sl = 0; sl < S; ++sl) {
                if (taken[sl]) continue;
                const int key = (sl / 8 / T) * 8 + sl % 8;
                const auto& u = use[(size_t)key * kBanks + bank];
                const int c = (u.empty() || std::find(u.begin(), u.end(), word) != u.end()) ? 0 : (int)u.size();
                if (c < best_cost) {
                    best_cost = c;
                    best = sl;
                }
            }
            taken[best] = 1;
            out[best] = r;
            const int key = (best / 8 / T) * 8 + best % 8;
            auto& u = use[(size_t)key * kBanks + bank];
            if (std::find(u.begin(), u.end(), word) == u.end()) {
                if (u.empty()) touched.push_back(key * kBanks + (int)bank);
                u.push_back(word);
            }
        }
        std::copy(out.begin(), out.end(), rec);
    }
};

// Order of the records within each adjacency row (unit format) for the
// split sweep's gathers.  A wave sweeps one replica: lane l takes the groups
// of 4 positions (cs >> 2) + l, + 64, ... of a class (positions outside the
// class idle), and the gather of record slot s of the lane's k-th position
// (k = p mod 4) is one byte read by every lane at once, each half-wave of 32
// lanes in its own LDS pass; two lanes of a half-wave that read different
// words of one bank serialise.  Rows are walked in position order; each
// record takes, among its row's free slots, the one whose gather already
// reads its word or reads the fewest other words in its bank (first such
// slot).  As with the cooperative order, any order gives the same sums.
struct SplitOrder {
    static constexpr int kBanks = 32, kCap = 32;  // a half-wave reads at most 32 words per gather
    int slots;                                    // row slots per k (max degree + 1)
    std::vector<uint8_t> cnt;                     // (k, slot, bank) -> words so far
    std::vector<uint16_t> words;                  // (k, slot, bank) -> those words
    std::vector<int32_t> touched;
    std::vector<int32_t> freeslots;
    int64_t group = -1;
    explicit SplitOrder(int max_deg)
        : slots(max_deg + 1), cnt((size_t)4 * slots * kBanks, 0), words((size_t)4 * slots * kBanks * kCap) {}
    // grp: (instance, class, iteration, half-wave); addr_mask keeps a record's
    // LDS byte address (position, and the sign-table offset when in use)
    void order(int64_t grp, int k, uint16_t* rec, int deg, uint32_t addr_mask) {
        if (grp != group) {
            for (int32_t t : touched) cnt[(size_t)t] = 0;
            touched.clear();
            group = grp;
        }
        freeslots.resize((size_t)deg);
        for (int q = 0; q < deg; ++q) freeslots[(size_t)q] = q;
        uint16_t out[kMaxDegree + 1];
        int nfree = deg;
        for (int j = 0; j < deg; ++j) {
            const uint32_t word = ((uint32_t)rec[j] & addr_mask) >> 2, bank = word % kBanks;
            int best = 0, best_cost = 1 << 30;
            for (int f = 0; f < nfree; ++f) {
                const int sl = freeslots[(size_t)f];
                const size_t e = ((size_t)k * slots + sl) * kBanks + bank;
                const int c = cnt[e];
                int cost = c;
                if (c > 0) {
                    const uint16_t* wv = &words[e * kCap];
                    for (int q = 0; q < c; ++q)
                        if (wv[q] == (uint16_t)word) {
                            cost = 0;
                            break;
                        }
                }
                if (cost < best_cost) {
                    best_cost = cost;
                    best = f;
                    if (cost == 0) break;
                }
            }
            const int sl = freeslots[(size_t)best];
            freeslots[(size_t)best] = freeslots[(size_t)(--nfree)];
            out[sl] = rec[j];
            const size_t e = ((size_t)k * slots + sl) * kBanks + bank;
            const int c = cnt[e];
            bool have = false;
            for (int q = 0; q < c && !have; ++q) have = words[e * kCap + q] == (uint16_t)word;
            if (!have && c < kCap) {
                if (c == 0) touched.push_back((int32_t)e);
                words[e * kCap + c] = (uint16_t)word;
                cnt[e] = (uint8_t)(c + 1);
            }
        }
        std::copy(out, out + deg, rec);
    }
};

// One problem's adjacency rows: rowinfo entries (row start in 16-byte units
// from the problem's first row << 9 | degree) and the padded rows.
struct RowBlock {
    std::vector<uint32_t> info, adj;
    int64_t len16 = 0;
    bool too_big = false;
};

// Table of the multispin lattice sweep (InstDesc::lsw_dim), appended to
// out; false when the instance does not qualify: a periodic L^D lattice
// (D = 2, 3) with even L, couplings +-1, no field, and positions that are the
// two site parities in index order (DSatur colours an even periodic lattice
// by parity, vertex 0 first; the check below confirms it for every
// position).  For each class c and window w of 32 class positions q (site
// x = 2k + par in row r = y + L z, q = r L/2 + k, par = (c + y + z) & 1):
// the edge signs per direction (+x, -x, +y, -y, +z, -z; bit = the coupling
// is negative) and the shape masks (par, k = 0, k = L/2 - 1, y = 0,
// y = L - 1, z = 0, z = L - 1).
bool lattice_sweep_qualifies(const Problem& P) {
    const int32_t D = P.lat_dim, L = P.lat_L;
    if (D < 2 || D > 3 || L < 4 || (L & 1)) return false;
    const int32_t H = L / 2;
    if (P.has_field || P.ncls != 2 || P.class_start[1] != P.n / 2) return false;
    for (int32_t wv : P.w)
        if (wv != 1 && wv != -1) return false;
    for (int32_t p = 0; p < P.n; ++p) {
        const int32_t o = P.pos2orig[p], x = o % L, r = o / L, y = r % L, z = r / L;
        if (p != ((x + y + z) & 1) * (P.n / 2) + r * H + x / 2) return false;
    }
    return true;
}

bool lattice_sweep_table(const Problem& P, std::vector<uint32_t>& out) {
    if (!lattice_sweep_qualifies(P)) return false;
    const int32_t D = P.lat_dim, L = P.lat_L;
    const int32_t H = L / 2;
    auto weight = [&](int32_t i, int32_t j) {
        for (int32_t q = P.rowptr[i]; q < P.rowptr[i + 1]; ++q)
            if (P.col[q] == j) return P.w[q];
        return 0;
    };
    const int32_t hc = P.n / 2, nwin = (hc + 31) / 32;
    const size_t base = out.size();
    out.resize(base + (size_t)2 * nwin * kLswGroup, 0u);
    for (int32_t c = 0; c < 2; ++c) {
        for (int32_t q = 0; q < hc; ++q) {
            const int32_t r = q / H, k = q % H, y = r % L, z = r / L, par = (c + y + z) & 1;
            const int32_t x = 2 * k + par, o = x + L * r;
            uint32_t* g = out.data() + base + ((size_t)c * nwin + q / 32) * kLswGroup;
            const uint32_t bit = 1u << (q % 32);
            const int32_t nb[6] = {(x + 1) % L + L * r, (x + L - 1) % L + L * r,
                                   x + L * ((y + 1) % L + L * z), x + L * ((y + L - 1) % L + L * z),
                                   x + L * (y + L * ((z + 1) % L)), x + L * (y + L * ((z + L - 1) % L))};
            for (int32_t j = 0; j < 2 * D; ++j)
                if (weight(o, nb[j]) < 0) g[j] |= bit;
            const bool m[7] = {par == 1, k == 0, k == H - 1, y == 0, y == L - 1, z == 0, z == L - 1};
            for (int32_t j = 0; j < 7; ++j)
                if (m[j]) g[8 + j] |= bit;
        }
    }
    return true;
}
}  // namespace

Status Engine::create(const std::vector<std::shared_ptr<const Problem>>& problems, const double* betas,
                      const int32_t* n_temps, uint64_t seed_, const ptx_engine_config& c) {
    cfg = c;
    seed = seed_;
    probs = problems;
    ninst = (int32_t)problems.size();
    if (ninst <= 0) return Status::err(PTX_ERR_INVALID, "no instances");
    if (ninst > (1 << 24)) return Status::err(PTX_ERR_UNSUPPORTED, "more than 2^24 instances");
    if (cfg.chains != 1 && cfg.chains != 2) return Status::err(PTX_ERR_INVALID, "chains must be 1 or 2");
    if (cfg.sweeps < 0) return Status::err(PTX_ERR_INVALID, "negative sweeps");
    const bool pilot = (cfg.flags & PTX_FLAG_PILOT) != 0;
    if (cfg.chains == 1 && !pilot && !(cfg.flags & PTX_FLAG_NO_EXCHANGE))
        return Status::err(PTX_ERR_INVALID, "a single chain needs the pilot exchange (PTX_FLAG_PILOT)");
    if (pilot && cfg.chains != 1) return Status::err(PTX_ERR_INVALID, "the pilot run has one chain");
    int device_count = 0;
    if (hipGetDeviceCount(&device_count) != hipSuccess || device_count == 0)
        return Status::err(PTX_ERR_DEVICE, "no HIP device available");
    PTX_TRY(hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate"));
    PTX_TRY(hip_check(hipHostMalloc((void**)&h_fault, 4 * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent),
                      "hipHostMalloc"));
    std::memset(h_fault, 0, 4 * sizeof(uint32_t));
    PTX_TRY(hip_check(hipHostGetDevicePointer((void**)&d_fault, h_fault, 0), "hipHostGetDevicePointer"));
    {
        int dev = 0, rate_khz = 0;
        PTX_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
        PTX_TRY(hip_check(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate"));
        // default bound 100 ms: a PT step of one replica's workgroup takes microseconds to a
        // millisecond, so a longer wait means a partner that is not running
        const int64_t us = cfg.coop_wait_us > 0 ? cfg.coop_wait_us : (cfg.coop_wait_us == 0 ? 100000 : 0);
        wait_ticks = (uint64_t)us * (uint64_t)(rate_khz > 0 ? rate_khz : 100000) / 1000u;
    }

    // ---- host layout ----
    desc.resize(ninst);
    ladders.resize(ninst);
    std::vector<int32_t> cls, hv, p2o;
    std::vector<uint32_t> rowinfo, thr;
    std::vector<uint32_t> adjw;  // padded rows as 32-bit words (both formats)
    std::vector<double> beta;
    std::vector<uint32_t> lat;
    std::vector<uint32_t> jneg;  // lattice edge signs, per instance
    std::vector<uint32_t> lsign; // edge signs of the multispin lattice sweep, per instance
    int32_t lsw_bytes = 0;       // LDS bytes per wave of the lattice sweep's largest instance
    struct LatShape {
        int32_t dim, L;
        const Problem* prob;
        int64_t off;
    };
    std::vector<LatShape> lat_shapes;
    int64_t beta_off = 0, rep_off = 0, spin_off = 0, hist_off = 0;
    const bool hist = (cfg.flags & PTX_FLAG_HISTOGRAM) != 0;
    int64_t t_off = 0;
    std::unordered_map<const Problem*, int32_t> first_copy;  // problem -> its first entry
    int32_t deg_max = 0;
    for (const auto& pp : problems) deg_max = std::max(deg_max, pp->max_deg);
    // record format: 16-bit when every instance has |w| = 1 and n <= 32768
    bool unit_ok = true;
    for (const auto& pp : problems) {
        if (pp->max_deg > kMaxDegree)
            return Status::err(PTX_ERR_UNSUPPORTED, "a vertex has more than 511 neighbours");
        if (pp->n > 32736) unit_ok = false;  // the sentinel position 32 * nw needs 15 bits
        for (int32_t wv : pp->w) unit_ok = unit_ok && (wv == 1 || wv == -1);
    }
    fmt = unit_ok ? kFmtUnit : kFmtWide;
    // instances the split sweep takes through the multispin lattice path
    // (it reads their state bits, never their spin bytes)
    std::vector<char> lsw_q((size_t)ninst);
    int64_t lsw_need = 0;
    for (int32_t i = 0; i < ninst; ++i) {
        const Problem& P = *problems[i];
        lsw_q[(size_t)i] = lattice_sweep_qualifies(P) ? 1 : 0;
        if (lsw_q[(size_t)i]) lsw_need = std::max<int64_t>(lsw_need, lsw_lds_bytes((P.n + 31) / 32, P.n, P.lat_dim));
    }
    // instances the split sweep may take replica-sliced (k_sweep_rs, 32
    // replicas per wave; kept when the engine runs the split kernels)
    std::vector<char> rs_q((size_t)ninst);
    for (int32_t i = 0; i < ninst; ++i) {
        const Problem& P = *problems[i];
        // auto: only instances whose colour classes average at least 512
        // positions (128 groups of 4), so a workgroup's 256 lanes find work in
        // most of a class's passes (measured: G(5000, 0.004), 625 per class,
        // sweeps 1.4x faster; G(1000, 0.02), 125 per class, 1.6x slower), and
        // with 32 replicas or more (full blocks share each gather 32 ways)
        const bool big = cfg.sliced_sweep > 0 || ((int64_t)P.n >= 512ll * std::max(P.ncls, 1) &&
                                                  (int64_t)n_temps[i] * cfg.chains >= 32);
        rs_q[(size_t)i] = cfg.sliced_sweep >= 0 && cfg.mode != PTX_MODE_COOP && cfg.mode != PTX_MODE_FUSED &&
                                  fmt == kFmtUnit && !lsw_q[(size_t)i] && !P.has_field && P.max_deg <= 63 && big &&
                                  sliced_lds_bytes((P.n + 31) / 32, P.fmax) <= 64 * 1024
                              ? 1
                              : 0;
    }
    // the split sweep's waves per workgroup (as set below)
    int64_t reps0 = 0;
    for (int32_t i = 0; i < ninst; ++i) reps0 += (int64_t)n_temps[i] * cfg.chains;
    const int64_t Wsweep = cfg.waves_per_group > 0 ? std::min(cfg.waves_per_group, 8) : (reps0 < 4 * 256 ? 1 : 4);
    // sign table of the split sweep: unit records and a negative coupling in
    // an instance the sweep gathers for, every such instance's table within
    // kSignTableMax bytes, and the sweep's waves' tables (or lattice states)
    // and threshold rows within 64 KiB
    tab = 0;
    if (fmt == kFmtUnit) {
        bool any_neg = false;
        int64_t tmax = 0, fmx = 1;
        for (int32_t i = 0; i < ninst; ++i) {
            const Problem& P = *problems[i];
            fmx = std::max<int64_t>(fmx, P.fmax);
            if (lsw_q[(size_t)i] || rs_q[(size_t)i]) continue;
            for (int32_t wv : P.w) any_neg = any_neg || wv < 0;
            tmax = std::max<int64_t>(tmax, sign_table_bytes((P.n + 31) / 32));
        }
        tab = any_neg && tmax <= kSignTableMax &&
                      Wsweep * (align_up(std::max(tmax, lsw_need), 16) + 4 * align_up(fmx, 4)) <= 64 * 1024
                  ? 1
                  : 0;
    }
    // Adjacency rows of every distinct problem, by position, each padded to
    // 16 bytes (at least one chunk); padding records name the sentinel
    // position 32 * nw (the wide format masks padding by weight 0, so there
    // any in-range position will do).  Built on worker threads, one problem at
    // a time each: ordering the records for the sweeps' gathers is most of the
    // host work of setting up a large batch.
    std::vector<int32_t> block_of((size_t)ninst);
    std::vector<const Problem*> block_prob;
    std::vector<char> block_rs;  // the problem's sweeps are replica-sliced: any record order will do
    {
        std::unordered_map<const Problem*, int32_t> bidx;
        for (int32_t i = 0; i < ninst; ++i) {
            auto it = bidx.find(problems[i].get());
            if (it == bidx.end()) {
                it = bidx.emplace(problems[i].get(), (int32_t)block_prob.size()).first;
                block_prob.push_back(problems[i].get());
                block_rs.push_back(rs_q[(size_t)i]);
            }
            block_of[(size_t)i] = it->second;
        }
    }
    std::vector<RowBlock> blocks(block_prob.size());
    // engines that run the cooperative kernel with one position per lane from
    // the start order each row's records for its gathers, the others for the
    // split sweep's (any order is exact)
    const bool coop_rows = fmt == kFmtUnit && cfg.spins_per_lane != 4 &&
                           (cfg.mode == PTX_MODE_COOP || (cfg.mode == PTX_MODE_AUTO && ninst == 1));
    const int bank_threads = kWave * (cfg.waves_per_group == 16 ? 16 : cfg.waves_per_group >= 8 ? 8 : 4);
    auto build_block = [&](size_t bi, BankOrder& bank_order, SplitOrder* split_order) {
        const Problem& P = *block_prob[bi];
        RowBlock& out = blocks[bi];
        const int32_t nw = (P.n + 31) / 32;
        const int32_t kbit = tab ? sign_table_bit(nw) : 15;
        const uint32_t sentinel = sentinel_pos(nw);
        out.info.reserve((size_t)P.n);
        int64_t cur16 = 0;
        int32_t cls_of = 0;
        std::vector<uint16_t> rec;
        for (int32_t p = 0; p < P.n; ++p) {
            const int32_t o = P.pos2orig[p];
            const int32_t dg = P.rowptr[o + 1] - P.rowptr[o];
            if (cur16 >= (1ll << 23)) {
                out.too_big = true;
                return;
            }
            out.info.push_back((uint32_t)(cur16 << 9) | (uint32_t)dg);
            if (fmt == kFmtUnit) {
                const int32_t chunks = std::max(1, (dg + 7) / 8);
                rec.assign((size_t)chunks * 8, (uint16_t)sentinel);
                for (int32_t j = 0; j < dg; ++j) {
                    const int32_t q = P.rowptr[o] + j;
                    rec[j] = (uint16_t)((uint32_t)P.orig2pos[P.col[q]] | (P.w[q] < 0 ? 1u << kbit : 0u));
                }
                while (p >= P.class_start[cls_of + 1]) ++cls_of;
                const int32_t c0 = P.class_start[cls_of];
                if (coop_rows) {
                    bank_order.order(((int64_t)bi << 12) | cls_of, p - c0, P.class_start[cls_of + 1] - c0, rec.data(),
                                     dg, (1u << kbit) - 1u);
                } else if (split_order && !block_rs[bi]) {
                    const int32_t gi = (p >> 2) - (c0 >> 2);  // group of 4 positions within the class
                    const int64_t grp = (((int64_t)bi << 12 | cls_of) << 20) | (int64_t)(gi >> 5);  // half-wave
                    split_order->order(grp, p & 3, rec.data(), dg, tab ? (2u << kbit) - 1u : (1u << kbit) - 1u);
                }
                for (int32_t c = 0; c < chunks * 4; ++c)
                    out.adj.push_back((uint32_t)rec[2 * c] | ((uint32_t)rec[2 * c + 1] << 16));
                cur16 += chunks;
            } else {
                const int32_t chunks = std::max(1, (dg + 3) / 4);
                for (int32_t j = 0; j < chunks * 4; ++j) {
                    if (j < dg) {
                        const int32_t q = P.rowptr[o] + j;
                        out.adj.push_back((uint32_t)P.orig2pos[P.col[q]] | ((uint32_t)(uint16_t)(int16_t)P.w[q] << 16));
                    } else {
                        out.adj.push_back(sentinel);  // weight 0: contributes nothing
                    }
                }
                cur16 += chunks;
            }
        }
        out.len16 = cur16;
    };
    {
        const bool split_rows = fmt == kFmtUnit && !coop_rows;
        const unsigned hw = std::thread::hardware_concurrency();
        const size_t nth = std::min<size_t>(blocks.size(), std::max(1u, std::min(16u, hw)));
        auto worker = [&](std::atomic<size_t>* next) {
            BankOrder bank_order(bank_threads);
            std::unique_ptr<SplitOrder> so(split_rows ? new SplitOrder(deg_max) : nullptr);
            for (size_t bi; (bi = next->fetch_add(1)) < blocks.size();) build_block(bi, bank_order, so.get());
        };
        std::atomic<size_t> next{0};
        if (nth <= 1) {
            worker(&next);
        } else {
            std::vector<std::thread> pool;
            for (size_t t = 0; t < nth; ++t) pool.emplace_back(worker, &next);
            for (auto& th : pool) th.join();
        }
        for (const RowBlock& blk : blocks)
            if (blk.too_big) return Status::err(PTX_ERR_UNSUPPORTED, "instance adjacency above 128 MiB");
    }
    for (int32_t i = 0; i < ninst; ++i) {
        const Problem& P = *problems[i];
        const int32_t nt = n_temps[i];
        if (nt < 1 || nt > 256) return Status::err(PTX_ERR_UNSUPPORTED, "temperatures per chain must be in [1, 256]");
        ladders[i].assign(betas + t_off, betas + t_off + nt);
        t_off += nt;
        InstDesc& d = desc[i];
        d.n = P.n;
        d.nw = (P.n + 31) / 32;
        d.nt = nt;
        d.nrep = nt * cfg.chains;
        d.ncls = P.ncls;
        d.cls0 = P.class_start[1] - P.class_start[0];
        d.fmax = P.fmax;
        d.has_field = P.has_field ? 1 : 0;
        d.neg = std::any_of(P.w.begin(), P.w.end(), [](int32_t w) { return w < 0; }) ? 1 : 0;
        d.kbit = tab ? sign_table_bit(d.nw) : 15;
        d.max_deg = P.max_deg;
        d.deg_mean = (int32_t)(P.nnz() / std::max(P.n, 1));
        int32_t nc = 0;
        if (cfg.chains == 2 && (cfg.flags & PTX_FLAG_ICM)) {
            nc = cfg.nc >= 0 ? std::min(cfg.nc, nt) : icm_count(ladders[i].data(), nt, P.char_scale);
        }
        d.nc = nc;
        if (nc > 0) coop_icm = true;
        d.inst_id = (uint32_t)i;
        // copies of one problem in the batch (entries holding the same
        // Problem) share its read-only tables: rows, adjacency, class bounds,
        // field and position order, lattice edge signs and sweep table, and the
        // thresholds when the ladder is the same too
        const auto seen = first_copy.find(&P);
        const int32_t s_i = seen == first_copy.end() ? -1 : seen->second;
        const InstDesc* s = s_i >= 0 ? &desc[(size_t)s_i] : nullptr;
        if (!s) first_copy.emplace(&P, i);
        if (s) {
            d.pos_off = s->pos_off;
            d.adj_off16 = s->adj_off16;
            d.adj_len16 = s->adj_len16;
            d.cls_off = s->cls_off;
        } else {
            // this problem's rows, built before the loop
            RowBlock& blk = blocks[(size_t)block_of[(size_t)i]];
            d.pos_off = (int64_t)rowinfo.size();
            d.adj_off16 = (int64_t)adjw.size() / 4;
            rowinfo.insert(rowinfo.end(), blk.info.begin(), blk.info.end());
            adjw.insert(adjw.end(), blk.adj.begin(), blk.adj.end());
            const int64_t cur16 = blk.len16;
            std::vector<uint32_t>().swap(blk.info);
            std::vector<uint32_t>().swap(blk.adj);
            d.adj_len16 = cur16;
            d.cls_off = (int64_t)cls.size();
            cls.insert(cls.end(), P.class_start.begin(), P.class_start.end());
            for (int32_t p = 0; p < P.n; ++p) {
                hv.push_back(P.h[P.pos2orig[p]]);
                p2o.push_back(P.pos2orig[p]);
            }
        }
        d.thr_off = (int64_t)thr.size();
        if (s && ladders[(size_t)s_i] == ladders[(size_t)i]) {
            d.thr_off = s->thr_off;
        } else if (P.fmax > 0) {
            std::vector<uint32_t> t((size_t)nt * P.fmax);
            threshold_table(ladders[i].data(), nt, P.fmax, t.data());
            thr.insert(thr.end(), t.begin(), t.end());
        }
        d.beta_off = beta_off;
        for (int32_t ch = 0; ch < cfg.chains; ++ch) beta.insert(beta.end(), ladders[i].begin(), ladders[i].end());
        beta_off += d.nrep;
        d.rep_off = rep_off;
        rep_off += d.nrep;
        d.spin_off = spin_off;
        spin_off += (int64_t)d.nrep * d.nw;
        // lattice tables, shared by every instance of one lattice shape
        // (the colouring, and so the position order, depends on the shape only)
        d.lat_dim = d.lat_L = d.lat_R = d.lat_words = 0;
        d.lat_off = 0;
        d.jn_off = -1;
        if (s && nc > 0 && s->nc > 0) {
            d.lat_dim = s->lat_dim;
            d.lat_L = s->lat_L;
            d.lat_R = s->lat_R;
            d.lat_words = s->lat_words;
            d.lat_off = s->lat_off;
            d.jn_off = s->jn_off;
        } else if (nc > 0 && P.lat_dim > 0) {
            const int32_t L = P.lat_L, R = (L + 31) / 32;
            int32_t rows = 1;
            for (int32_t k = 1; k < P.lat_dim; ++k) rows *= L;
            int64_t off = -1;
            for (const auto& t : lat_shapes)
                if (t.dim == P.lat_dim && t.L == L && t.prob->pos2orig == P.pos2orig) off = t.off;
            if (off < 0) {
                off = (int64_t)lat.size();
                const int32_t words = R * rows;
                std::vector<uint16_t> wpos((size_t)words * 32, 0);
                for (int32_t w = 0; w < words; ++w) {
                    const int32_t row = w / R, k = w % R;
                    for (int32_t b = 0; b < 32 && 32 * k + b < L; ++b)
                        wpos[(size_t)w * 32 + b] = (uint16_t)P.orig2pos[row * L + 32 * k + b];
                }
                for (size_t q = 0; q < wpos.size(); q += 2) lat.push_back((uint32_t)wpos[q] | ((uint32_t)wpos[q + 1] << 16));
                for (int32_t p = 0; p < P.n; ++p) {
                    const int32_t o = P.pos2orig[p], x = o % L, row = o / L;
                    lat.push_back((uint32_t)((row * R + x / 32) << 5) | (uint32_t)(x % 32));
                }
                while (lat.size() % 4) lat.push_back(0u);  // 16-byte aligned tables
                lat_shapes.push_back({P.lat_dim, L, problems[i].get(), off});
            }
            d.lat_dim = P.lat_dim;
            d.lat_L = L;
            d.lat_R = R;
            d.lat_words = R * rows;
            d.lat_off = off;
            lat_words_max = std::max(lat_words_max, d.lat_words);
            // edge signs for the cut energy on bit sets (L >= 4: the lattice-
            // order S1 fits the position-list area), unit couplings, no field
            if (L >= 4 && fmt == kFmtUnit && !P.has_field) {
                d.jn_off = (int64_t)jneg.size();
                auto weight = [&](int32_t i, int32_t j) {
                    for (int32_t q = P.rowptr[i]; q < P.rowptr[i + 1]; ++q)
                        if (P.col[q] == j) return P.w[q];
                    return 0;
                };
                int64_t stride = 1;
                for (int32_t ax = 0; ax < P.lat_dim; ++ax, stride *= L) {
                    for (int32_t row = 0; row < rows; ++row) {
                        for (int32_t k = 0; k < R; ++k) {
                            uint32_t word = 0;
                            for (int32_t b = 0; b < 32 && 32 * k + b < L; ++b) {
                                const int32_t o = row * L + 32 * k + b;
                                const int32_t c = (int32_t)((o / stride) % L);
                                const int32_t j = o - (int32_t)(c * stride) + (int32_t)(((c + 1) % L) * stride);
                                if (weight(o, j) < 0) word |= 1u << b;
                            }
                            jneg.push_back(word);
                        }
                    }
                }
            }
        }
        d.lsw_dim = d.lsw_L = 0;
        d.lsw_off = (int64_t)lsign.size();
        if (s) {
            d.lsw_dim = s->lsw_dim;
            d.lsw_L = s->lsw_L;
            d.lsw_off = s->lsw_off;
        } else if (lattice_sweep_table(P, lsign)) {
            d.lsw_dim = P.lat_dim;
            d.lsw_L = P.lat_L;
            lsw_bytes = std::max(lsw_bytes, lsw_lds_bytes(d.nw, d.n, d.lsw_dim));
        }
        d.e_base = (int32_t)(-P.abs_sum);
        d.hist_bins = hist ? (int32_t)(P.abs_sum + 1) : 0;
        d.hist_off = hist_off;
        if (hist) hist_off += (int64_t)d.nrep * d.hist_bins;
        nrep_max = std::max(nrep_max, d.nrep);
        nw_max = std::max(nw_max, d.nw);
        n_max = std::max(n_max, d.n);
        fmax_max = std::max(fmax_max, std::max(1, d.fmax));
        updates_per_step += (int64_t)d.nrep * P.n * cfg.sweeps;
        lds_bytes_per_step += (int64_t)d.nrep * cfg.sweeps * (4 * P.nnz() + 6ll * P.n);
    }
    total_rep = rep_off;

    // ---- launch shape ----
    W = cfg.waves_per_group;
    if (W <= 0) W = (total_rep < 4 * 256) ? 1 : 4;  // spread few replicas over CUs, group many
    if (W > 8) W = 8;                                 // split kernels take at most 8 waves
    if (W != 1 && W != 2 && W != 4 && W != 8) return Status::err(PTX_ERR_INVALID, "waves_per_group must be 1,2,4,8,16");
    if (cfg.sliced_sweep != 0 && cfg.sliced_sweep != -1 && cfg.sliced_sweep != 1 && cfg.sliced_sweep != 2 &&
        cfg.sliced_sweep != 4 && cfg.sliced_sweep != 8)
        return Status::err(PTX_ERR_INVALID, "sliced_sweep must be 0, -1, 1, 2, 4 or 8");
    Wi = cfg.icm_waves;
    if (Wi != 0 && Wi != 1 && Wi != 2 && Wi != 4) return Status::err(PTX_ERR_INVALID, "icm_waves must be 0, 1, 2 or 4");
    // per-wave LDS of the split sweep: the spin bytes (or sign tables) of the
    // instances it gathers for and the multispin lattices' states; k_init
    // expands every instance's spin bytes
    int64_t gather_need = 0;
    for (int32_t i = 0; i < ninst; ++i)
        if (!lsw_q[(size_t)i] && !rs_q[(size_t)i])
            gather_need = std::max<int64_t>(gather_need, tab ? sign_table_bytes(desc[i].nw) : spin_bytes(desc[i].nw));
    spin_stride = (int32_t)align_up(std::max<int64_t>(gather_need, lsw_bytes), 16);
    init_stride = (int32_t)align_up(std::max<int64_t>(spin_stride, spin_bytes(nw_max)), 16);
    thr_stride = (int32_t)align_up(std::max(fmax_max, 1), 4);  // at least one word: the sweep may read entry 0
    const int64_t lds = (int64_t)W * (spin_stride + 4ll * thr_stride);
    if (lds > 64 * 1024 || (int64_t)W * init_stride > 64 * 1024)
        return Status::err(PTX_ERR_UNSUPPORTED, "sweep workgroup exceeds 64 KiB of LDS; use fewer waves");
    // work items: largest instances first so the tail of the grid holds the small ones
    std::vector<int32_t> order(ninst);
    for (int32_t i = 0; i < ninst; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
        return (int64_t)desc[x].n * desc[x].max_deg > (int64_t)desc[y].n * desc[y].max_deg;
    });
    size_order = order;
    active.assign((size_t)ninst, 1);

    // ---- execution mode ----
    spin_words = spin_off;
    mode = PTX_MODE_SPLIT;
    const bool observe = (cfg.flags & PTX_FLAG_OBSERVE) != 0;
    if (observe && (cfg.mode == PTX_MODE_COOP || cfg.mode == PTX_MODE_FUSED))
        return Status::err(PTX_ERR_UNSUPPORTED, "observables (PTX_FLAG_OBSERVE) are taken by the split kernels only");
    if (observe) {
        // split kernels: the state is in device memory after every step
    } else if (cfg.mode == PTX_MODE_COOP) {
        spl = cfg.spins_per_lane == 4 ? 4 : 1;
        // 16 waves only for one position per lane and no cluster moves
        coop_W = cfg.waves_per_group == 16 && spl == 1 && !coop_icm ? 16
                 : cfg.waves_per_group >= 8                        ? 8
                                                                   : 4;
        coop_K = cfg.steps_per_launch > 0 ? cfg.steps_per_launch : 256;
        PTX_TRY(setup_coop(true));
        mode = PTX_MODE_COOP;
    } else if (cfg.mode == PTX_MODE_AUTO && ninst == 1) {
        coop_W = 4;
        spl = cfg.spins_per_lane == 4 ? 4 : 1;
        coop_K = cfg.steps_per_launch > 0 ? cfg.steps_per_launch : 256;
        PTX_TRY(setup_coop(false));
        if (coop_capacity >= total_rep) mode = PTX_MODE_COOP;
    }
    Wf = coop_W;
    if (mode == PTX_MODE_COOP) K = coop_K;
    // The fused kernel runs when asked for (or chosen by the autotuner): on
    // the batches measured here the split kernels, which spread every
    // instance's replicas and cluster moves over the whole GPU, are faster.
    if (mode == PTX_MODE_SPLIT && cfg.mode == PTX_MODE_FUSED) {
        int Wpref = cfg.waves_per_group > 0 ? cfg.waves_per_group : (nrep_max >= 32 ? 16 : 8);
        if (Wpref != 4 && Wpref != 8 && Wpref != 16) Wpref = Wpref < 6 ? 4 : (Wpref < 12 ? 8 : 16);
        // LDS bytes of the largest instance for W waves, adjacency in LDS or not
        auto need = [&](int w, int csr_lds) {
            int mx = 0;
            for (int32_t i = 0; i < ninst; ++i)
                mx = std::max(mx, fused_layout(desc[i].n, (int)desc[i].adj_len16, desc[i].nrep, desc[i].nt,
                                               desc[i].fmax, w, csr_lds, desc[i].nc > 0 ? 1 : 0)
                                      .total);
            return mx;
        };
        const int lds_limit = 160 * 1024;
        const int place = cfg.csr_placement;
        // an explicit waves_per_group is kept; otherwise fewer waves may make an instance fit
        std::vector<int> ws;
        for (int w = Wpref; w >= 4; w /= 2) {
            ws.push_back(w);
            if (cfg.waves_per_group > 0) break;
        }
        bool placed = false;
        for (int csr_lds = 1; csr_lds >= 0 && !placed; --csr_lds) {
            if ((csr_lds && place == PTX_CSR_GLOBAL) || (!csr_lds && place == PTX_CSR_LDS)) continue;
            for (int w : ws) {
                const int b = need(w, csr_lds);
                if (b <= lds_limit) {
                    mode = PTX_MODE_FUSED;
                    Wf = w;
                    fused_lds = b;
                    fused_global = csr_lds ? 0 : 1;
                    placed = true;
                    break;
                }
            }
        }
        if (!placed && cfg.mode == PTX_MODE_FUSED)
            return Status::err(PTX_ERR_UNSUPPORTED, "instance does not fit the fused kernel's LDS (" +
                                                        std::to_string(need(ws.back(), place == PTX_CSR_LDS ? 1 : 0)) +
                                                        " bytes at " + std::to_string(ws.back()) +
                                                        " waves); use split mode");
        K = cfg.steps_per_launch > 0 ? cfg.steps_per_launch : 16;
    }
    // replica-sliced sweeps in split mode; k_sweep_rs's LDS fits the largest
    rs_stride = 0;
    for (int32_t i = 0; i < ninst; ++i) {
        InstDesc& d = desc[(size_t)i];
        d.rs = mode == PTX_MODE_SPLIT && rs_q[(size_t)i] ? 1 : 0;
        if (d.rs) rs_stride = std::max(rs_stride, (int32_t)sliced_lds_bytes(d.nw, d.fmax));
    }
    // waves per 32 replicas: as configured, else 4 (measured on G(5000,
    // 0.004), whose classes take 157 groups of 4: 4 waves 0.72 ms per sweep of
    // 256 instances, 3 waves 0.81 ms though fewer lanes idle, 2 waves 1.15 ms,
    // 8 waves 1.03 ms; more waves hide more latency until most lanes idle)
    rs_W = cfg.sliced_sweep > 0 ? cfg.sliced_sweep : 4;
    // ---- device buffers ----
    PTX_TRY(d_desc.upload(desc, stream));
    PTX_TRY(d_inst_step.alloc((size_t)ninst));
    if (mode == PTX_MODE_SPLIT && cfg.small_batch_coop && !observe &&
        (cfg.mode == PTX_MODE_AUTO || cfg.mode == PTX_MODE_SPLIT)) {
        coop_W = 4;
        spl = 1;
        coop_K = 256;
        PTX_TRY(setup_coop(false));
    }
    if (coop_capacity > 0) {
        PTX_TRY(d_pub_spins.alloc((size_t)(2 * spin_words)));
        PTX_TRY(d_pub_E.alloc((size_t)(2 * total_rep)));
        PTX_TRY(d_err.alloc(1));
    }
    last_mode = mode;
    PTX_TRY(d_rowinfo.upload(rowinfo, stream));
    PTX_TRY(d_adj.upload(adjw, stream));
    PTX_TRY(d_cls.upload(cls, stream));
    PTX_TRY(d_h.upload(hv, stream));
    PTX_TRY(d_p2o.upload(p2o, stream));
    if (lat.empty()) lat.assign(4, 0u);
    PTX_TRY(d_lat.upload(lat, stream));
    if (jneg.empty()) jneg.assign(4, 0u);
    PTX_TRY(d_jneg.upload(jneg, stream));
    if (lsign.empty()) lsign.assign(kLswGroup, 0u);
    PTX_TRY(d_lsign.upload(lsign, stream));
    PTX_TRY(d_thr.upload(thr, stream));
    PTX_TRY(d_beta.upload(beta, stream));
    PTX_TRY(build_lists());
    PTX_TRY(d_spins.alloc((size_t)spin_off));
    PTX_TRY(d_best_spins.alloc((size_t)spin_off));
    PTX_TRY(d_E.alloc((size_t)total_rep));
    PTX_TRY(d_bestE.alloc((size_t)total_rep));
    PTX_TRY(d_best_sweep.alloc((size_t)total_rep));
    PTX_TRY(d_slot2rep.alloc((size_t)total_rep));
    PTX_TRY(d_rep2slot.alloc((size_t)total_rep));
    PTX_TRY(d_flips.alloc((size_t)total_rep));
    PTX_TRY(d_att.alloc((size_t)total_rep));
    PTX_TRY(d_succ.alloc((size_t)total_rep));
    PTX_TRY(d_swaps.alloc((size_t)ninst));
    PTX_TRY(d_hist.alloc((size_t)std::max<int64_t>(hist_off, 1)));
    PTX_TRY(d_obs_i.alloc(observe ? (size_t)total_rep * 4 : 1));
    PTX_TRY(d_obs_d.alloc(observe ? (size_t)total_rep * 2 : 1));
    PTX_TRY(hip_check(hipStreamSynchronize(stream), "upload sync"));
    PTX_TRY(launch_init(0));
    return sync();
}

static DevArgs build_args(Engine& e) {
    DevArgs a{};
    a.desc = e.d_desc.p;
    a.rowinfo = e.d_rowinfo.p;
    a.adj = reinterpret_cast<const uint8_t*>(e.d_adj.p);
    a.fmt = e.fmt;
    a.cls = e.d_cls.p;
    a.hfield = e.d_h.p;
    a.p2o = e.d_p2o.p;
    a.thr = e.d_thr.p;
    a.beta = e.d_beta.p;
    a.spins = e.d_spins.p;
    a.best_spins = e.d_best_spins.p;
    a.E = e.d_E.p;
    a.bestE = e.d_bestE.p;
    a.best_sweep = e.d_best_sweep.p;
    a.flips = e.d_flips.p;
    a.slot2rep = e.d_slot2rep.p;
    a.rep2slot = e.d_rep2slot.p;
    a.swap_att = e.d_att.p;
    a.swap_succ = e.d_succ.p;
    a.swaps = e.d_swaps.p;
    a.hist = e.d_hist.p;
    a.items = e.d_items.p;
    a.pairs = e.d_pairs.p;
    a.order = e.d_order.p;
    a.lat = e.d_lat.p;
    a.jneg = e.d_jneg.p;
    a.lsign = e.d_lsign.p;
    a.phase = e.phase_on ? e.d_phase.p : nullptr;
    a.fault = e.d_fault;
    a.obs_i = reinterpret_cast<long long*>(e.d_obs_i.p);
    a.obs_d = e.d_obs_d.p;
    a.inst_step = e.d_inst_step.p;
    a.seed = e.seed;
    a.tag_base = (e.cfg.flags & PTX_FLAG_PILOT) ? kTagPilot : 0u;
    a.sweeps = e.cfg.sweeps;
    a.spin_stride = e.spin_stride;
    a.tab = e.tab;
    a.thr_stride = e.thr_stride;
    a.pilot = (e.cfg.flags & PTX_FLAG_PILOT) ? 1 : 0;
    a.no_exchange = (e.cfg.flags & PTX_FLAG_NO_EXCHANGE) ? 1 : 0;
    a.hist_on = (e.cfg.flags & PTX_FLAG_HISTOGRAM) ? 1 : 0;
    a.hist_start = e.cfg.hist_start;
    return a;
}

Status Engine::launch_init(int mode) {
    DevArgs a = build_args(*this);
    a.spin_stride = init_stride;  // every instance's spin bytes, lattices included
    const size_t lds = (size_t)W * init_stride;
    const dim3 grid((unsigned)items.size()), block((unsigned)(W * kWave));
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, block, lds, stream, a, mode); };
    if (fmt == kFmtUnit) {
        switch (W) {
            case 1: go(k_init<kFmtUnit, 1>); break;
            case 2: go(k_init<kFmtUnit, 2>); break;
            case 4: go(k_init<kFmtUnit, 4>); break;
            default: go(k_init<kFmtUnit, 8>); break;
        }
    } else {
        switch (W) {
            case 1: go(k_init<kFmtWide, 1>); break;
            case 2: go(k_init<kFmtWide, 2>); break;
            case 4: go(k_init<kFmtWide, 4>); break;
            default: go(k_init<kFmtWide, 8>); break;
        }
    }
    return hip_check(hipGetLastError(), "k_init launch");
}

Status Engine::launch_step() {
    DevArgs a = build_args(*this);
    auto timed = [&](int kind, auto&& fn) -> Status {
        if (!timing) {
            fn();
            return hip_check(hipGetLastError(), "kernel launch");
        }
        Timed t;
        t.kind = kind;
        PTX_TRY(hip_check(hipEventCreate(&t.a), "hipEventCreate"));
        PTX_TRY(hip_check(hipEventCreate(&t.b), "hipEventCreate"));
        PTX_TRY(hip_check(hipEventRecord(t.a, stream), "hipEventRecord"));
        fn();
        PTX_TRY(hip_check(hipGetLastError(), "kernel launch"));
        PTX_TRY(hip_check(hipEventRecord(t.b, stream), "hipEventRecord"));
        pending.push_back(t);
        return Status::ok();
    };
    if (cfg.sweeps > 0 && (!rs_items.empty() || !sweep_items.empty())) {
        // one timed phase: the replica-sliced instances, then the others
        PTX_TRY(timed(0, [&] {
            if (!rs_items.empty()) {
                DevArgs ar = a;
                ar.items = d_rs_items.p;
                const int Wr = rs_W;
                auto go = [&](auto kern) {
                    hipLaunchKernelGGL(kern, dim3((unsigned)rs_items.size()), dim3((unsigned)(Wr * kWave)),
                                       (size_t)rs_stride, stream, ar);
                };
                switch (Wr) {
                    case 1: go(k_sweep_rs<1>); break;
                    case 2: go(k_sweep_rs<2>); break;
                    case 8: go(k_sweep_rs<8>); break;
                    default: go(k_sweep_rs<4>); break;
                }
            }
            if (sweep_items.empty()) return;
            DevArgs as = a;
            as.items = d_sweep_items.p;
            const size_t lds = (size_t)W * (spin_stride + 4 * (size_t)thr_stride);
            const dim3 grid((unsigned)sweep_items.size()), block((unsigned)(W * kWave));
            auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, block, lds, stream, as); };
            if (fmt == kFmtUnit) {
                switch (W) {
                    case 1: go(k_sweep<kFmtUnit, 1>); break;
                    case 2: go(k_sweep<kFmtUnit, 2>); break;
                    case 4: go(k_sweep<kFmtUnit, 4>); break;
                    default: go(k_sweep<kFmtUnit, 8>); break;
                }
            } else {
                switch (W) {
                    case 1: go(k_sweep<kFmtWide, 1>); break;
                    case 2: go(k_sweep<kFmtWide, 2>); break;
                    case 4: go(k_sweep<kFmtWide, 4>); break;
                    default: go(k_sweep<kFmtWide, 8>); break;
                }
            }
        }));
    }
    {
        const size_t lds = (size_t)nrep_max * 48 + 64;
        PTX_TRY(timed(1, [&] {
            hipLaunchKernelGGL(k_exchange, dim3((unsigned)inst_list.size()), dim3(kWave), lds, stream, a, nrep_max);
        }));
    }
    if (!pairs.empty()) {
        const int wi = icm_waves_now();
        const size_t lds = 8 * kIcmRedSlots + (size_t)(nw_max + 1) * 16 + 16 + (32 * (size_t)nw_max + 16) +
                           8 * (size_t)lat_words_max;
        const dim3 grid((unsigned)pairs.size()), block((unsigned)(wi * kWave));
        PTX_TRY(timed(2, [&] {
            auto go = [&](auto kern) { hipLaunchKernelGGL(kern, grid, block, lds, stream, a, nw_max, lat_words_max); };
            if (fmt == kFmtUnit) {
                switch (wi) {
                    case 1: go(k_icm<kFmtUnit, 1>); break;
                    case 2: go(k_icm<kFmtUnit, 2>); break;
                    default: go(k_icm<kFmtUnit, 4>); break;
                }
            } else {
                switch (wi) {
                    case 1: go(k_icm<kFmtWide, 1>); break;
                    case 2: go(k_icm<kFmtWide, 2>); break;
                    default: go(k_icm<kFmtWide, 4>); break;
                }
            }
        }));
    }
    const int na = (int)inst_list.size();
    if (cfg.flags & PTX_FLAG_OBSERVE) {
        hipLaunchKernelGGL(k_observe, dim3((unsigned)na), dim3(kWave), 0, stream, a);
        PTX_TRY(hip_check(hipGetLastError(), "k_observe"));
    }
    hipLaunchKernelGGL(k_step_advance, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, stream, d_order.p, na,
                       d_inst_step.p);
    return hip_check(hipGetLastError(), "k_step_advance");
}

Status Engine::launch_fused(int nsteps) {
    DevArgs a = build_args(*this);
    const dim3 grid((unsigned)inst_list.size()), block((unsigned)(Wf * kWave));
    const size_t lds = (size_t)fused_lds;
    auto go = [&](auto kern) -> Status {
        if (lds > 64 * 1024)
            PTX_TRY(hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds),
                              "hipFuncSetAttribute"));
        Timed t;
        if (timing) {
            t.kind = 3;
            PTX_TRY(hip_check(hipEventCreate(&t.a), "hipEventCreate"));
            PTX_TRY(hip_check(hipEventCreate(&t.b), "hipEventCreate"));
            PTX_TRY(hip_check(hipEventRecord(t.a, stream), "hipEventRecord"));
        }
        hipLaunchKernelGGL(kern, grid, block, lds, stream, a, nsteps);
        PTX_TRY(hip_check(hipGetLastError(), "k_fused launch"));
        if (timing) {
            PTX_TRY(hip_check(hipEventRecord(t.b, stream), "hipEventRecord"));
            pending.push_back(t);
        }
        return Status::ok();
    };
    if (fused_global) {
        if (fmt == kFmtUnit) {
            switch (Wf) {
                case 4: return go(k_fused<kFmtUnit, 4, 1>);
                case 8: return go(k_fused<kFmtUnit, 8, 1>);
                default: return go(k_fused<kFmtUnit, 16, 1>);
            }
        }
        switch (Wf) {
            case 4: return go(k_fused<kFmtWide, 4, 1>);
            case 8: return go(k_fused<kFmtWide, 8, 1>);
            default: return go(k_fused<kFmtWide, 16, 1>);
        }
    }
    if (fmt == kFmtUnit) {
        switch (Wf) {
            case 4: return go(k_fused<kFmtUnit, 4, 0>);
            case 8: return go(k_fused<kFmtUnit, 8, 0>);
            default: return go(k_fused<kFmtUnit, 16, 0>);
        }
    }
    switch (Wf) {
        case 4: return go(k_fused<kFmtWide, 4, 0>);
        case 8: return go(k_fused<kFmtWide, 8, 0>);
        default: return go(k_fused<kFmtWide, 16, 0>);
    }
}

// The cooperative kernel of the engine's format, width and positions per
// lane, with the adjacency in LDS (g = 0) or read through L2 (g = 1: four
// waves, one position per lane), and with cluster moves compiled in when
// any instance has cluster-move slots.
const void* Engine::coop_kernel(int g) const {
    auto pick = [&](auto fmt_tag, auto ic_tag) -> const void* {
        constexpr int F = decltype(fmt_tag)::value, I = decltype(ic_tag)::value;
        if (g) return reinterpret_cast<const void*>(k_coop<F, 4, 1, 1, I>);
        if (coop_W == 16) return reinterpret_cast<const void*>(k_coop<F, 16, 1, 0, 0>);
        if (coop_W == 8) return spl == 4 ? reinterpret_cast<const void*>(k_coop<F, 8, 4, 0, I>)
                                         : reinterpret_cast<const void*>(k_coop<F, 8, 1, 0, I>);
        return spl == 4 ? reinterpret_cast<const void*>(k_coop<F, 4, 4, 0, I>)
                        : reinterpret_cast<const void*>(k_coop<F, 4, 1, 0, I>);
    };
    using U = std::integral_constant<int, kFmtUnit>;
    using Wd = std::integral_constant<int, kFmtWide>;
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    if (fmt == kFmtUnit) return coop_icm ? pick(U{}, I1{}) : pick(U{}, I0{});
    return coop_icm ? pick(Wd{}, I1{}) : pick(Wd{}, I0{});
}

// Sizes the cooperative launches, for the adjacency in LDS and (with four
// waves and one position per lane) read through L2: the LDS layout of the
// largest instance (with the lattice tables and the next sweep's random
// words when every instance still fits 160 KiB, tables first to go) and the
// number of workgroups the device keeps resident at once.  required: an
// error when the whole batch fits neither form; coop_capacity = 0 says no
// cooperative launch is possible.
Status Engine::setup_coop(bool required) {
    coop_capacity = 0;
    int dev = 0, ncu = 0, coopok = 0;
    PTX_TRY(hip_check(hipGetDevice(&dev), "hipGetDevice"));
    PTX_TRY(hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev), "CU count"));
    PTX_TRY(hip_check(hipDeviceGetAttribute(&coopok, hipDeviceAttributeCooperativeLaunch, dev), "coop attr"));
    const int lds_limit = 160 * 1024;
    for (int g = 0; g < 2; ++g) {
        CoopForm& cf = coop_form[g];
        cf = CoopForm{};
        if (!coopok || (g == 1 && (coop_W != 4 || spl != 1))) continue;
        auto max_total = [&](int pre, int stage, int f0) {
            int mx = 0;
            for (const InstDesc& d : desc)
                mx = std::max(mx, coop_layout(d.n, (int)d.adj_len16, d.ncls, d.nrep, d.nt, d.fmax, pre, d.lat_words,
                                              stage, g, f0 ? d.cls0 : 0)
                                      .total);
            return mx;
        };
        // class-0 fields drawn during the step wait: kernels without cluster
        // moves (a replica's state then does not change between sweeps)
        cf.pre_f0 = coop_icm ? 0 : 1;
        cf.pre_rnd = 1;
        cf.lat_stage = 1;
        if (max_total(1, 1, cf.pre_f0) > lds_limit) cf.pre_f0 = 0;
        if (max_total(1, 1, 0) > lds_limit) cf.lat_stage = 0;
        if (max_total(1, 0, 0) > lds_limit) cf.pre_rnd = 0;
        if (!cf.pre_rnd) cf.pre_f0 = 0;
        cf.lds = max_total(cf.pre_rnd, cf.lat_stage, cf.pre_f0);
        if (cf.lds > lds_limit) continue;
        const void* f = coop_kernel(g);
        if (cf.lds > 64 * 1024)
            PTX_TRY(hip_check(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, cf.lds),
                              "hipFuncSetAttribute"));
        int bpc = 0;
        PTX_TRY(hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, f, coop_W * kWave, (size_t)cf.lds),
                          "occupancy"));
        cf.capacity = (int64_t)bpc * ncu;
        coop_capacity = std::max(coop_capacity, cf.capacity);
    }
    if (required && coop_capacity < total_rep)
        return Status::err(PTX_ERR_UNSUPPORTED,
                           "cooperative mode needs every replica of the batch resident at once (" +
                               std::to_string(total_rep) + " workgroups, room for " + std::to_string(coop_capacity) +
                               ") and each instance within 160 KiB of LDS");
    return Status::ok();
}

Status Engine::launch_coop(int nsteps) {
    // the adjacency in LDS while the active replicas fit that form, else through L2
    const int g = (int64_t)coop_items.size() <= coop_form[0].capacity ? 0 : 1;
    const CoopForm& cf = coop_form[g];
    coop_lds = cf.lds;
    coop_last_form = g;
    DevArgs a = build_args(*this);
    CoopArgs c{};
    c.pub_spins = d_pub_spins.p;
    c.pub_E = d_pub_E.p;
    c.items = d_coop_items.p;
    c.err = d_err.p;
    c.fault = d_fault;
    c.wait_ticks = wait_ticks;
    c.pub_words = spin_words;
    c.pub_reps = (int32_t)total_rep;
    c.pre_rnd = cf.pre_rnd;
    c.lat_stage = cf.lat_stage;
    c.pre_f0 = cf.pre_f0;
    const dim3 grid((unsigned)coop_items.size()), block((unsigned)(coop_W * kWave));
    const unsigned lds = (unsigned)cf.lds;
    const void* f = coop_kernel(g);
    void* args[] = {&a, &c, &nsteps};
    Timed t;
    if (timing) {
        t.kind = 3;
        PTX_TRY(hip_check(hipEventCreate(&t.a), "hipEventCreate"));
        PTX_TRY(hip_check(hipEventCreate(&t.b), "hipEventCreate"));
        PTX_TRY(hip_check(hipEventRecord(t.a, stream), "hipEventRecord"));
    }
    PTX_TRY(hip_check(hipLaunchCooperativeKernel(f, grid, block, args, lds, stream), "k_coop launch"));
    if (timing) {
        PTX_TRY(hip_check(hipEventRecord(t.b, stream), "hipEventRecord"));
        pending.push_back(t);
    }
    return Status::ok();
}

Status Engine::step(int64_t n) {
    if (n < 0) return Status::err(PTX_ERR_INVALID, "negative step count");
    PTX_TRY(check_fault());  // no more launches after a reported fault
    if (inst_list.empty()) return Status::ok();  // every instance retired
    if (steps_done + n > 0xFFFFFFFFll) return Status::err(PTX_ERR_UNSUPPORTED, "more than 2^32 PT steps");
    // cooperative launches: the cooperative mode, or a split-mode engine
    // with small_batch_coop whose active replicas fit one launch
    if (mode == PTX_MODE_COOP || (mode == PTX_MODE_SPLIT && cfg.small_batch_coop && coop_capacity > 0 &&
                                  active_reps <= coop_capacity)) {
        last_mode = PTX_MODE_COOP;
        while (n > 0) {
            const int k = (int)std::min<int64_t>(n, coop_K);
            PTX_TRY(check_fault());
            PTX_TRY(launch_coop(k));
            steps_done += k;
            n -= k;
        }
        return Status::ok();
    }
    last_mode = mode;
    if (mode == PTX_MODE_FUSED) {
        const int64_t G = cfg.use_graph > 0 ? cfg.use_graph : 0;
        while (n > 0) {
            if (G > 0 && graph_exec == nullptr && steps_done > 0 && n >= G * K) PTX_TRY(build_graph());
            if (graph_exec != nullptr && n >= G * K) {
                Timed t;
                if (timing) {
                    t.kind = 3;
                    t.count = (int)G;
                    PTX_TRY(hip_check(hipEventCreate(&t.a), "hipEventCreate"));
                    PTX_TRY(hip_check(hipEventCreate(&t.b), "hipEventCreate"));
                    PTX_TRY(hip_check(hipEventRecord(t.a, stream), "hipEventRecord"));
                }
                PTX_TRY(hip_check(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch"));
                if (timing) {
                    PTX_TRY(hip_check(hipEventRecord(t.b, stream), "hipEventRecord"));
                    pending.push_back(t);
                }
                steps_done += G * K;
                n -= G * K;
                continue;
            }
            const int k = (int)std::min<int64_t>(n, K);
            PTX_TRY(launch_fused(k));  // the first launch also sets the kernel's LDS attribute
            steps_done += k;
            n -= k;
        }
        return Status::ok();
    }
    // split mode: with use_graph = G (and timing off, which needs per-kernel
    // events) G steps of launches are captured once and replayed
    const int64_t G = (cfg.use_graph > 0 && !timing) ? cfg.use_graph : 0;
    while (n > 0) {
        if (G > 0 && n >= G) {
            if (graph_exec == nullptr) PTX_TRY(build_graph());
            PTX_TRY(hip_check(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch"));
            steps_done += G;
            n -= G;
            continue;
        }
        PTX_TRY(launch_step());
        ++steps_done;
        --n;
    }
    return Status::ok();
}

// Uploads a work list, growing its buffer only when it is too small (lists
// shrink as instances retire).
template <class T>
static Status put_list(DevBuf<T>& buf, const std::vector<T>& v, hipStream_t s) {
    const size_t need = std::max<size_t>(v.size(), 1);
    if (buf.p == nullptr || buf.n < need) PTX_TRY(buf.alloc(need));
    if (v.empty()) return Status::ok();
    return buf.upload(v.data(), v.size(), 0, s);
}

// Work lists of the active instances.  MI355X hands consecutive workgroups
// of a launch to its 8 XCDs in turn, each with its own L2.  With many
// instances, instance k of the size order goes to XCD k % 8 (round robin
// keeps the XCDs' work even) and all of its sweep and cluster-move items run
// there, so its adjacency is cached in one L2 instead of eight: launch slot b
// holds item b / 8 of XCD b % 8's list, short lists padded with idle items
// (instance -1).  inst_list (one workgroup per instance: exchange, fused
// kernel) lists the active instances in size order.
Status Engine::build_lists() {
    items.clear();
    pairs.clear();
    inst_list.clear();
    std::vector<int32_t> act;
    for (int32_t i : size_order)
        if (active[(size_t)i]) act.push_back(i);
    inst_list = act;
    updates_per_step = 0;
    lds_bytes_per_step = 0;
    for (int32_t i : act) {
        const Problem& P = *probs[(size_t)i];
        updates_per_step += (int64_t)desc[i].nrep * P.n * cfg.sweeps;
        lds_bytes_per_step += (int64_t)desc[i].nrep * cfg.sweeps * (4 * P.nnz() + 6ll * P.n);
    }
    sweep_items.clear();
    rs_items.clear();
    if (act.size() >= 8) {
        std::vector<std::vector<int2>> ix(8), px(8), sx(8), rx(8);
        for (size_t k = 0; k < act.size(); ++k) {
            const int32_t i = act[k];
            const int x = (int)(k % 8);
            for (int32_t r0 = 0; r0 < desc[i].nrep; r0 += W) ix[x].push_back(make_int2(i, r0));
            for (int32_t t = desc[i].nt - 1; t >= desc[i].nt - desc[i].nc; --t) px[x].push_back(make_int2(i, t));
            if (desc[i].rs) {
                for (int32_t r0 = 0; r0 < desc[i].nrep; r0 += 32) rx[x].push_back(make_int2(i, r0));
            } else {
                for (int32_t r0 = 0; r0 < desc[i].nrep; r0 += W) sx[x].push_back(make_int2(i, r0));
            }
        }
        auto arrange = [](const std::vector<std::vector<int2>>& per, std::vector<int2>& out) {
            size_t len = 0;
            for (const auto& v : per) len = std::max(len, v.size());
            out.assign(len * 8, make_int2(-1, 0));
            for (int x = 0; x < 8; ++x)
                for (size_t j = 0; j < per[x].size(); ++j) out[j * 8 + x] = per[x][j];
        };
        arrange(ix, items);
        arrange(px, pairs);
        arrange(sx, sweep_items);
        arrange(rx, rs_items);
        // an empty kind needs no idle items
        if (std::all_of(sx.begin(), sx.end(), [](const std::vector<int2>& v) { return v.empty(); }))
            sweep_items.clear();
        if (std::all_of(rx.begin(), rx.end(), [](const std::vector<int2>& v) { return v.empty(); }))
            rs_items.clear();
    } else {
        for (int32_t i : act)
            for (int32_t r0 = 0; r0 < desc[i].nrep; r0 += W) items.push_back(make_int2(i, r0));
        for (int32_t i : act) {
            if (desc[i].rs) {
                for (int32_t r0 = 0; r0 < desc[i].nrep; r0 += 32) rs_items.push_back(make_int2(i, r0));
            } else {
                for (int32_t r0 = 0; r0 < desc[i].nrep; r0 += W) sweep_items.push_back(make_int2(i, r0));
            }
        }
        for (int32_t i : act)
            for (int32_t t = desc[i].nt - 1; t >= desc[i].nt - desc[i].nc; --t) pairs.push_back(make_int2(i, t));
    }
    coop_items.clear();
    active_reps = 0;
    for (int32_t i : act) {
        active_reps += desc[i].nrep;
        for (int32_t r = 0; r < desc[i].nrep; ++r) coop_items.push_back(make_int2(i, r));
    }
    PTX_TRY(put_list(d_items, items, stream));
    PTX_TRY(put_list(d_sweep_items, sweep_items, stream));
    PTX_TRY(put_list(d_rs_items, rs_items, stream));
    PTX_TRY(put_list(d_pairs, pairs, stream));
    PTX_TRY(put_list(d_order, inst_list, stream));
    PTX_TRY(put_list(d_coop_items, coop_items, stream));
    return hip_check(hipStreamSynchronize(stream), "work lists");
}

Status Engine::set_active(const uint8_t* flags) {
    PTX_TRY(sync());
    if (graph_exec) {  // the graph holds the old grid sizes
        (void)hipGraphExecDestroy(graph_exec);
        (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }
    for (int32_t i = 0; i < ninst; ++i) active[(size_t)i] = flags[i] ? 1 : 0;
    return build_lists();
}

// Captures use_graph fused launches of K steps, or use_graph split-mode PT
// steps (sweep, exchange, cluster moves, step advance), on the engine
// stream.  No launch here is a memset or a host copy, so the graph holds
// kernel nodes only; every kernel reads its step from device memory, so a
// replay continues where the last one ended.
Status Engine::build_graph() {
    const bool was_timing = timing;
    timing = false;
    PTX_TRY(hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture"));
    Status st = Status::ok();
    for (int32_t g = 0; g < cfg.use_graph && st; ++g) st = mode == PTX_MODE_FUSED ? launch_fused(K) : launch_step();
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &gr);
    timing = was_timing;
    if (!st) {
        if (gr) (void)hipGraphDestroy(gr);
        return st;
    }
    PTX_TRY(hip_check(e, "hipStreamEndCapture"));
    graph = gr;
    return hip_check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

Status Engine::sync() {
    PTX_TRY(hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize"));
    return check_fault();
}

Status Engine::collect_timing() {
    PTX_TRY(sync());
    for (auto& t : pending) {
        float m = 0.f;
        PTX_TRY(hip_check(hipEventElapsedTime(&m, t.a, t.b), "hipEventElapsedTime"));
        ms[t.kind] += m;
        launches[t.kind] += t.count;
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    pending.clear();
    return Status::ok();
}

Status Engine::timing_get(ptx_engine_timing* out) {
    PTX_TRY(collect_timing());
    out->sweep_ms = ms[0];
    out->exchange_ms = ms[1];
    out->icm_ms = ms[2];
    out->fused_ms = ms[3];
    out->sweep_launches = launches[0];
    out->exchange_launches = launches[1];
    out->icm_launches = launches[2];
    out->fused_launches = launches[3];
    return Status::ok();
}

Status Engine::phase_profile(int on, double* us) {
    if (on >= 0) {
        if (!d_phase.p) PTX_TRY(d_phase.alloc(4));
        PTX_TRY(sync());
        PTX_TRY(hip_check(hipMemset(d_phase.p, 0, 4 * sizeof(unsigned long long)), "hipMemset"));
        phase_on = on != 0;
    }
    if (us) {
        unsigned long long t[4] = {0, 0, 0, 0};
        if (d_phase.p) PTX_TRY(d_phase.download(t, 4, 0, stream));
        int rate_khz = 0;
        PTX_TRY(hip_check(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, 0), "wall clock rate"));
        for (int k = 0; k < 4; ++k) us[k] = rate_khz > 0 ? (double)t[k] * 1e3 / rate_khz : 0.0;
    }
    return Status::ok();
}

static Status check_inst(const Engine& e, int32_t inst) {
    if (inst < 0 || inst >= e.ninst) return Status::err(PTX_ERR_INVALID, "instance index out of range");
    return Status::ok();
}

Status Engine::read_spins(int32_t inst, uint8_t* out) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    const InstDesc& d = desc[inst];
    const Problem& P = *probs[inst];
    std::vector<uint32_t> w((size_t)d.nrep * d.nw);
    PTX_TRY(d_spins.download(w.data(), w.size(), (size_t)d.spin_off, stream));
    for (int32_t r = 0; r < d.nrep; ++r)
        for (int32_t p = 0; p < d.n; ++p)
            out[(size_t)r * d.n + P.pos2orig[p]] = (uint8_t)((w[(size_t)r * d.nw + (p >> 5)] >> (p & 31)) & 1u);
    return Status::ok();
}

Status Engine::write_spins(int32_t inst, const uint8_t* in) {
    PTX_TRY(check_inst(*this, inst));
    if (!active[(size_t)inst])  // its energies are recomputed by the active work list
        return Status::err(PTX_ERR_INVALID, "instance is retired; reactivate it before writing its states");
    const InstDesc& d = desc[inst];
    const Problem& P = *probs[inst];
    std::vector<uint32_t> w((size_t)d.nrep * d.nw, 0u);
    for (int32_t r = 0; r < d.nrep; ++r)
        for (int32_t p = 0; p < d.n; ++p)
            if (in[(size_t)r * d.n + P.pos2orig[p]]) w[(size_t)r * d.nw + (p >> 5)] |= 1u << (p & 31);
    PTX_TRY(d_spins.upload(w.data(), w.size(), (size_t)d.spin_off, stream));
    PTX_TRY(launch_init(1));
    return sync();
}

Status Engine::read_energies(int32_t inst, int64_t* out) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    const InstDesc& d = desc[inst];
    std::vector<int32_t> e(d.nrep);
    PTX_TRY(d_E.download(e.data(), e.size(), (size_t)d.rep_off, stream));
    for (int32_t r = 0; r < d.nrep; ++r) out[r] = e[r];
    return Status::ok();
}

Status Engine::read_perm(int32_t inst, int32_t* out) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    const InstDesc& d = desc[inst];
    return d_slot2rep.download(out, (size_t)d.nrep, (size_t)d.rep_off, stream);
}

Status Engine::read_swap_stats(int32_t inst, uint32_t* att, uint32_t* succ, uint64_t* swaps) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    const InstDesc& d = desc[inst];
    PTX_TRY(d_att.download(att, (size_t)d.nrep, (size_t)d.rep_off, stream));
    PTX_TRY(d_succ.download(succ, (size_t)d.nrep, (size_t)d.rep_off, stream));
    unsigned long long s = 0;
    PTX_TRY(d_swaps.download(&s, 1, (size_t)inst, stream));
    *swaps = s;
    return Status::ok();
}

Status Engine::read_flips(int32_t inst, uint64_t* out) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    const InstDesc& d = desc[inst];
    std::vector<unsigned long long> f(d.nrep);
    PTX_TRY(d_flips.download(f.data(), f.size(), (size_t)d.rep_off, stream));
    for (int32_t r = 0; r < d.nrep; ++r) out[r] = f[r];
    return Status::ok();
}

Status Engine::read_best(int32_t inst, int64_t* be, int64_t* bs, uint8_t* st) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    const InstDesc& d = desc[inst];
    const Problem& P = *probs[inst];
    std::vector<int32_t> e(d.nrep), s(d.nrep);
    PTX_TRY(d_bestE.download(e.data(), e.size(), (size_t)d.rep_off, stream));
    PTX_TRY(d_best_sweep.download(s.data(), s.size(), (size_t)d.rep_off, stream));
    for (int32_t r = 0; r < d.nrep; ++r) {
        be[r] = e[r] == 0x7FFFFFFF ? INT64_MAX : (int64_t)e[r];
        bs[r] = s[r];
    }
    if (st) {
        std::vector<uint32_t> w((size_t)d.nrep * d.nw);
        PTX_TRY(d_best_spins.download(w.data(), w.size(), (size_t)d.spin_off, stream));
        for (int32_t r = 0; r < d.nrep; ++r)
            for (int32_t p = 0; p < d.n; ++p)
                st[(size_t)r * d.n + P.pos2orig[p]] = (uint8_t)((w[(size_t)r * d.nw + (p >> 5)] >> (p & 31)) & 1u);
    }
    return Status::ok();
}

Status Engine::read_histograms(int32_t inst, uint32_t* out) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    const InstDesc& d = desc[inst];
    if (!(cfg.flags & PTX_FLAG_HISTOGRAM)) return Status::err(PTX_ERR_INVALID, "engine built without histograms");
    return d_hist.download(out, (size_t)d.nrep * d.hist_bins, (size_t)d.hist_off, stream);
}

Status Engine::read_observables(int32_t inst, ptx_slot_observables* out) {
    PTX_TRY(check_inst(*this, inst));
    PTX_TRY(sync());
    if (!(cfg.flags & PTX_FLAG_OBSERVE)) return Status::err(PTX_ERR_INVALID, "engine built without observables");
    const InstDesc& d = desc[inst];
    std::vector<unsigned long long> vi((size_t)d.nrep * 4);
    std::vector<double> vd((size_t)d.nrep * 2);
    PTX_TRY(d_obs_i.download(vi.data(), vi.size(), (size_t)d.rep_off * 4, stream));
    PTX_TRY(d_obs_d.download(vd.data(), vd.size(), (size_t)d.rep_off * 2, stream));
    for (int32_t s = 0; s < d.nrep; ++s) {
        out[s].samples = (int64_t)vi[4 * s];
        out[s].sum_e = (int64_t)vi[4 * s + 1];
        out[s].sum_up = (int64_t)vi[4 * s + 2];
        out[s].sum_abs_m = (int64_t)vi[4 * s + 3];
        out[s].sum_e2 = vd[2 * s];
        out[s].sum_up2 = vd[2 * s + 1];
    }
    return Status::ok();
}

// ---- checkpoint ----
// Blob: a header that identifies the engine's shape (what load_state checks
// before it writes anything), then every array a PT step changes, in a fixed
// order.  The random streams are indexed by (seed, instance, step), so a
// restored engine continues exactly as the saved one would have.
namespace {
constexpr uint32_t kStateMagic = 0x53585450u;  // "PTXS"
constexpr uint32_t kStateVersion = 1;
struct StateHeader {
    uint32_t magic, version;
    uint64_t shape;       // FNV-1a of the instance descriptors, ladders, flags, sweeps and seed
    int64_t ninst, total_rep, spin_words, hist_len, observe;
    int64_t steps_done;
};
struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void add(const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
};
}  // namespace

static uint64_t shape_hash(const Engine& e) {
    Fnv f;
    for (const InstDesc& d : e.desc) {
        const int32_t v[9] = {d.n, d.nw, d.nt, d.nrep, d.ncls, d.fmax, d.nc, d.hist_bins, d.e_base};
        f.add(v, sizeof v);
    }
    for (const auto& l : e.ladders) f.add(l.data(), l.size() * sizeof(double));
    const uint32_t fl = e.cfg.flags;
    const int32_t sw = e.cfg.sweeps, ch = e.cfg.chains;
    f.add(&fl, sizeof fl);
    f.add(&sw, sizeof sw);
    f.add(&ch, sizeof ch);
    f.add(&e.seed, sizeof e.seed);
    f.add(&e.cfg.hist_start, sizeof e.cfg.hist_start);
    return f.h;
}

int64_t Engine::state_bytes() const {
    const int64_t R = total_rep, I = ninst;
    const int64_t hist_len = (int64_t)d_hist.n, obs = (cfg.flags & PTX_FLAG_OBSERVE) ? 1 : 0;
    return (int64_t)sizeof(StateHeader) + I * 4 /*step*/ + I /*active*/ + 2 * spin_words * 4 + 5 * R * 4 +
           R * 8 /*flips*/ + I * 8 /*swaps*/ + 2 * R * 4 /*att, succ*/ + hist_len * 4 + obs * R * (4 * 8 + 2 * 8);
}

Status Engine::save_state(void* buf, int64_t bytes) {
    if (bytes < state_bytes()) return Status::err(PTX_ERR_INVALID, "state buffer too small");
    PTX_TRY(sync());
    uint8_t* o = static_cast<uint8_t*>(buf);
    StateHeader h{kStateMagic, kStateVersion, shape_hash(*this), ninst, total_rep, spin_words, (int64_t)d_hist.n,
                  (cfg.flags & PTX_FLAG_OBSERVE) ? 1 : 0, steps_done};
    std::memcpy(o, &h, sizeof h);
    o += sizeof h;
    auto take = [&](auto& dev, size_t count) -> Status {
        using T = std::remove_pointer_t<decltype(dev.p)>;
        PTX_TRY(dev.download(reinterpret_cast<T*>(o), count, 0, stream));
        o += count * sizeof(T);
        return Status::ok();
    };
    const size_t R = (size_t)total_rep, I = (size_t)ninst, S = (size_t)spin_words;
    PTX_TRY(take(d_inst_step, I));
    std::memcpy(o, active.data(), I);
    o += I;
    PTX_TRY(take(d_spins, S));
    PTX_TRY(take(d_best_spins, S));
    PTX_TRY(take(d_E, R));
    PTX_TRY(take(d_bestE, R));
    PTX_TRY(take(d_best_sweep, R));
    PTX_TRY(take(d_slot2rep, R));
    PTX_TRY(take(d_rep2slot, R));
    PTX_TRY(take(d_flips, R));
    PTX_TRY(take(d_swaps, I));
    PTX_TRY(take(d_att, R));
    PTX_TRY(take(d_succ, R));
    PTX_TRY(take(d_hist, d_hist.n));
    if (h.observe) {
        PTX_TRY(take(d_obs_i, 4 * R));
        PTX_TRY(take(d_obs_d, 2 * R));
    }
    return Status::ok();
}

Status Engine::load_state(const void* buf, int64_t bytes) {
    if (bytes < (int64_t)sizeof(StateHeader)) return Status::err(PTX_ERR_INVALID, "state buffer too small");
    const uint8_t* o = static_cast<const uint8_t*>(buf);
    StateHeader h;
    std::memcpy(&h, o, sizeof h);
    o += sizeof h;
    if (h.magic != kStateMagic || h.version != kStateVersion)
        return Status::err(PTX_ERR_INVALID, "not an engine state (magic or version)");
    if (h.shape != shape_hash(*this) || h.ninst != ninst || h.total_rep != total_rep || h.spin_words != spin_words ||
        h.hist_len != (int64_t)d_hist.n || h.observe != ((cfg.flags & PTX_FLAG_OBSERVE) ? 1 : 0))
        return Status::err(PTX_ERR_INVALID, "the state was saved by an engine of another shape (instances, "
                                            "ladders, flags, sweeps or seed differ)");
    if (bytes < state_bytes()) return Status::err(PTX_ERR_INVALID, "state buffer too small");
    PTX_TRY(sync());
    auto put = [&](auto& dev, size_t count) -> Status {
        using T = std::remove_pointer_t<decltype(dev.p)>;
        PTX_TRY(dev.upload(reinterpret_cast<const T*>(o), count, 0, stream));
        o += count * sizeof(T);
        return Status::ok();
    };
    const size_t R = (size_t)total_rep, I = (size_t)ninst, S = (size_t)spin_words;
    PTX_TRY(put(d_inst_step, I));
    std::vector<uint8_t> act(o, o + I);
    o += I;
    PTX_TRY(put(d_spins, S));
    PTX_TRY(put(d_best_spins, S));
    PTX_TRY(put(d_E, R));
    PTX_TRY(put(d_bestE, R));
    PTX_TRY(put(d_best_sweep, R));
    PTX_TRY(put(d_slot2rep, R));
    PTX_TRY(put(d_rep2slot, R));
    PTX_TRY(put(d_flips, R));
    PTX_TRY(put(d_swaps, I));
    PTX_TRY(put(d_att, R));
    PTX_TRY(put(d_succ, R));
    PTX_TRY(put(d_hist, d_hist.n));
    if (h.observe) {
        PTX_TRY(put(d_obs_i, 4 * R));
        PTX_TRY(put(d_obs_d, 2 * R));
    }
    // the cooperative step's tagged energy slots of earlier launches could
    // match a step of the restored run: clear them
    if (d_pub_E.p) PTX_TRY(hip_check(hipMemsetAsync(d_pub_E.p, 0, d_pub_E.n * sizeof(unsigned long long), stream),
                                     "clear published energies"));
    steps_done = h.steps_done;
    PTX_TRY(hip_check(hipStreamSynchronize(stream), "load state"));
    return set_active(act.data());  // rebuilds the work lists (and drops a captured graph)
}

Status Engine::read_min_energies(int64_t* out) {
    PTX_TRY(sync());
    std::vector<int32_t> e((size_t)total_rep);
    PTX_TRY(d_bestE.download(e.data(), e.size(), 0, stream));
    for (int32_t i = 0; i < ninst; ++i) {
        int64_t m = INT64_MAX;
        for (int32_t r = 0; r < desc[i].nrep; ++r) {
            const int32_t v = e[(size_t)desc[i].rep_off + r];
            if (v != 0x7FFFFFFF) m = std::min<int64_t>(m, v);
        }
        out[i] = m;
    }
    return Status::ok();
}

// ---- self-tests ----
Status selftest_fp64(const double* x, int64_t n, double* eo, double* lo) {
    if (n <= 0) return Status::ok();
    DevBuf<double> dx, de, dl;
    PTX_TRY(dx.alloc((size_t)n));
    PTX_TRY(de.alloc((size_t)n));
    PTX_TRY(dl.alloc((size_t)n));
    PTX_TRY(hip_check(hipMemcpy(dx.p, x, n * sizeof(double), hipMemcpyHostToDevice), "copy"));
    hipLaunchKernelGGL(k_selftest_fp64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dx.p, n, de.p, dl.p);
    PTX_TRY(hip_check(hipGetLastError(), "k_selftest_fp64"));
    PTX_TRY(hip_check(hipMemcpy(eo, de.p, n * sizeof(double), hipMemcpyDeviceToHost), "copy"));
    return hip_check(hipMemcpy(lo, dl.p, n * sizeof(double), hipMemcpyDeviceToHost), "copy");
}

Status selftest_philox(const uint32_t* ctr, int64_t n, const uint32_t* key, uint32_t* out) {
    if (n <= 0) return Status::ok();
    DevBuf<uint32_t> dc, dout;
    PTX_TRY(dc.alloc((size_t)n * 4));
    PTX_TRY(dout.alloc((size_t)n * 4));
    PTX_TRY(hip_check(hipMemcpy(dc.p, ctr, n * 16, hipMemcpyHostToDevice), "copy"));
    hipLaunchKernelGGL(k_selftest_philox, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, dc.p, n, key[0], key[1],
                       dout.p);
    PTX_TRY(hip_check(hipGetLastError(), "k_selftest_philox"));
    return hip_check(hipMemcpy(out, dout.p, n * 16, hipMemcpyDeviceToHost), "copy");
}

Status selftest_copy(int64_t nbytes, int32_t reps, double* ms_per_copy) {
    const int64_t n16 = nbytes / 16;
    if (n16 <= 0 || reps <= 0) return Status::err(PTX_ERR_INVALID, "copy size and repetitions must be positive");
    DevBuf<uint32_t> a, b;
    PTX_TRY(a.alloc((size_t)n16 * 4));
    PTX_TRY(b.alloc((size_t)n16 * 4));
    hipEvent_t e0, e1;
    PTX_TRY(hip_check(hipEventCreate(&e0), "event"));
    PTX_TRY(hip_check(hipEventCreate(&e1), "event"));
    const unsigned grid = (unsigned)std::min<int64_t>((n16 + 255) / 256, 65536);
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, 0, reinterpret_cast<const uint4*>(a.p),
                       reinterpret_cast<uint4*>(b.p), n16);
    PTX_TRY(hip_check(hipEventRecord(e0, 0), "event"));
    for (int32_t r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, 0, reinterpret_cast<const uint4*>(a.p),
                           reinterpret_cast<uint4*>(b.p), n16);
    PTX_TRY(hip_check(hipEventRecord(e1, 0), "event"));
    PTX_TRY(hip_check(hipEventSynchronize(e1), "event"));
    float ms = 0;
    PTX_TRY(hip_check(hipEventElapsedTime(&ms, e0, e1), "event"));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_copy = ms / reps;
    return Status::ok();
}

Status selftest_lds(int32_t blocks, int32_t iters, double* gbs) {
    if (blocks <= 0 || iters <= 0) return Status::err(PTX_ERR_INVALID, "blocks and iterations must be positive");
    DevBuf<uint32_t> o;
    PTX_TRY(o.alloc((size_t)blocks));
    hipEvent_t e0, e1;
    PTX_TRY(hip_check(hipEventCreate(&e0), "event"));
    PTX_TRY(hip_check(hipEventCreate(&e1), "event"));
    hipLaunchKernelGGL(k_lds_read, dim3(blocks), dim3(256), 0, 0, o.p, iters, 0x9E3779B9u, 256);
    PTX_TRY(hip_check(hipEventRecord(e0, 0), "event"));
    hipLaunchKernelGGL(k_lds_read, dim3(blocks), dim3(256), 0, 0, o.p, iters, 0x9E3779B9u, 256);
    PTX_TRY(hip_check(hipEventRecord(e1, 0), "event"));
    PTX_TRY(hip_check(hipEventSynchronize(e1), "event"));
    float ms = 0;
    PTX_TRY(hip_check(hipEventElapsedTime(&ms, e0, e1), "event"));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbs = (double)blocks * 256.0 * 16.0 * (double)((iters + 3) / 4 * 4) / (ms * 1e-3) / 1e9;
    return Status::ok();
}

template struct DevBuf<uint32_t>;
template struct DevBuf<uint16_t>;
template struct DevBuf<unsigned long long>;
template struct DevBuf<int32_t>;
template struct DevBuf<double>;
template struct DevBuf<float>;
template struct DevBuf<uint8_t>;

}  // namespace ptx
