// engine.hpp -- host side of the engine: device buffers, instance layout,
// launch orchestration on one HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "../../include/ptx.h"
#include "problem.hpp"
#include "ptx_common.h"

#define PTX_TRY(expr)                  \
    do {                               \
        ::ptx::Status _s = (expr);     \
        if (!_s) return _s;            \
    } while (0)

namespace ptx {

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
        o.p = nullptr;
        o.n = 0;
    }
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    Status alloc(size_t count);
    Status upload(const std::vector<T>& v, hipStream_t s);
    Status upload(const T* v, size_t count, size_t offset, hipStream_t s);
    Status download(T* v, size_t count, size_t offset, hipStream_t s) const;
};

struct Engine {
    ptx_engine_config cfg{};
    uint64_t seed = 0;
    int32_t ninst = 0;
    std::vector<std::shared_ptr<const Problem>> probs;
    std::vector<InstDesc> desc;
    std::vector<std::vector<double>> ladders;  // per instance, nt entries
    int64_t total_rep = 0;
    int64_t steps_done = 0;
    int32_t W = 1;               // waves per sweep workgroup (split mode)
    int32_t Wi = 0;              // waves per ICM slot pair (split mode; 0 = by the number of pairs)
    int32_t mode = PTX_MODE_SPLIT;
    int32_t fmt = 0;             // adjacency record format of the fused kernel
    int32_t tab = 0;             // unit records carry the sign at the instance's kbit; k_sweep keeps the sign table
    int32_t Wf = 8;              // waves per workgroup (fused mode)
    int32_t K = 16;              // PT steps per fused launch
    int32_t fused_lds = 0;       // bytes of dynamic LDS per fused workgroup
    int32_t fused_global = 0;    // fused mode reads the adjacency from global memory
    int32_t nrep_max = 0, nw_max = 0, n_max = 0, fmax_max = 1;
    int32_t lat_words_max = 0;   // largest lattice-order bit set (words)
    int32_t spin_stride = 0, thr_stride = 1;
    int32_t init_stride = 0;  // k_init's LDS bytes per wave (spin bytes of the largest instance)
    int64_t updates_per_step = 0, lds_bytes_per_step = 0;
    std::vector<int2> items, pairs;
    // split sweeps: k_sweep's items (instances with rs = 0) and k_sweep_rs's
    // (instance, first of 32 replicas) items; items lists every instance (k_init)
    std::vector<int2> sweep_items, rs_items;
    int32_t rs_stride = 0;  // k_sweep_rs's LDS bytes (position words and threshold rows)
    int32_t rs_W = 4;       // k_sweep_rs's waves per workgroup (32 replicas)
    std::vector<int32_t> inst_list;  // active instances in size order (one workgroup each)
    std::vector<int32_t> size_order; // all instances, largest first
    std::vector<uint8_t> active;
    hipStream_t stream = nullptr;
    bool timing = false;
    struct Timed {
        hipEvent_t a, b;
        int kind;
        int count = 1;  // kernel launches between the two events
    };
    std::vector<Timed> pending;
    double ms[4] = {0, 0, 0, 0};
    int64_t launches[4] = {0, 0, 0, 0};

    DevBuf<InstDesc> d_desc;
    DevBuf<int32_t> d_cls, d_h, d_p2o;
    DevBuf<uint32_t> d_lat;      // lattice tables (wpos, obit) per lattice shape
    DevBuf<uint32_t> d_jneg;     // lattice edge signs per instance (cut energies on bit sets)
    DevBuf<uint32_t> d_lsign;    // edge signs of the multispin lattice sweep per instance (InstDesc::lsw_off)
    DevBuf<uint32_t> d_rowinfo, d_adj, d_thr;
    DevBuf<double> d_beta;
    DevBuf<uint32_t> d_spins, d_best_spins;
    DevBuf<int32_t> d_E, d_bestE, d_best_sweep, d_slot2rep, d_rep2slot;
    DevBuf<unsigned long long> d_flips, d_swaps;
    DevBuf<uint32_t> d_att, d_succ, d_hist;
    DevBuf<unsigned long long> d_obs_i;  // per slot: samples, sum E, sum up, sum |2 up - n| (PTX_FLAG_OBSERVE)
    DevBuf<double> d_obs_d;              // per slot: sum E^2, sum up^2
    DevBuf<int2> d_items, d_pairs, d_sweep_items, d_rs_items;
    DevBuf<int32_t> d_order;
    DevBuf<uint32_t> d_inst_step;
    DevBuf<unsigned long long> d_phase;
    // cooperative launches (cooperative mode, or small batches of a split-mode
    // engine with small_batch_coop): one workgroup per active replica
    DevBuf<uint32_t> d_pub_spins, d_err;
    DevBuf<unsigned long long> d_pub_E;  // tagged energies (step + 1) << 32 | E, two per replica
    DevBuf<int2> d_coop_items;
    std::vector<int2> coop_items;  // (instance, replica) per workgroup, active instances
    int64_t spin_words = 0;      // spin words of all instances
    int64_t active_reps = 0;     // replicas of the active instances
    int64_t coop_capacity = 0;   // co-resident cooperative workgroups (0: no cooperative launches)
    int32_t coop_W = 4;          // waves per cooperative workgroup
    int32_t coop_K = 256;        // PT steps per cooperative launch
    int32_t last_mode = PTX_MODE_SPLIT;  // mode of the latest launches
    int32_t spl = 1;             // positions per lane (cooperative mode)
    struct CoopForm {            // a cooperative launch form: adjacency in LDS (0) or read through L2 (1)
        int32_t lds = 0;         // dynamic LDS bytes per workgroup
        int32_t pre_rnd = 0;     // next sweep's random words drawn during the step wait
        int32_t lat_stage = 0;   // lattice tables copied into LDS
        int32_t pre_f0 = 0;      // next sweep's class-0 local fields drawn during the step wait
        int64_t capacity = 0;    // co-resident workgroups (0: form not available)
    };
    CoopForm coop_form[2];
    int32_t coop_lds = 0;        // LDS bytes of the latest cooperative launch
    int32_t coop_last_form = 0;  // form of the latest cooperative launch
    bool coop_icm = false;       // some instance has cluster-move slots (kernel with cluster moves)
    bool phase_on = false;
    // fused mode with use_graph = G > 0: G launches of K steps captured once
    // and replayed (the kernels read their step counter from device memory)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // fault words in pinned, device-mapped host memory, written by kernels
    // only on a fault and read by the host without a synchronisation:
    // [0] a cooperative step wait gave up, [1] histogram samples outside their table
    uint32_t* h_fault = nullptr;
    uint32_t* d_fault = nullptr;  // the same words as the kernels address them
    uint64_t wait_ticks = 0;      // bound of one cooperative step wait (wall-clock ticks)

    ~Engine();
    Status create(const std::vector<std::shared_ptr<const Problem>>& problems, const double* betas,
                  const int32_t* n_temps, uint64_t seed, const ptx_engine_config& c);
    Status step(int64_t n);
    Status sync();
    Status read_spins(int32_t inst, uint8_t* out);
    Status write_spins(int32_t inst, const uint8_t* in);
    Status read_energies(int32_t inst, int64_t* out);
    Status read_perm(int32_t inst, int32_t* out);
    Status read_swap_stats(int32_t inst, uint32_t* att, uint32_t* succ, uint64_t* swaps);
    Status read_flips(int32_t inst, uint64_t* out);
    Status read_best(int32_t inst, int64_t* be, int64_t* bs, uint8_t* st);
    Status read_histograms(int32_t inst, uint32_t* out);
    Status read_min_energies(int64_t* out);
    Status read_observables(int32_t inst, ptx_slot_observables* out);
    // checkpoint of the whole mutable state (see ptx_engine_save_state)
    int64_t state_bytes() const;
    Status save_state(void* buf, int64_t bytes);
    Status load_state(const void* buf, int64_t bytes);
    Status timing_get(ptx_engine_timing* out);
    Status phase_profile(int on, double* us);

  private:
    Status launch_init(int mode);
    Status launch_step();
    Status launch_fused(int nsteps);
    Status launch_coop(int nsteps);
    Status setup_coop(bool required);
    const void* coop_kernel(int g) const;
    Status build_graph();
    Status build_lists();

  public:
    Status set_active(const uint8_t* flags);
    // waves per ICM slot pair: the configured number, else 1.  Clusters grow
    // level by level and their frontiers are mostly narrower than one wave
    // (2D lattices: ~10 sites per level over ~30 levels), so more waves
    // only add barriers; the autotuner still tries 4.
    int icm_waves_now() const { return Wi > 0 ? Wi : 1; }

    Status collect_timing();
    // PTX_ERR_DEVICE once a kernel has reported a fault (sticky: the state is
    // no longer a valid trajectory)
    Status check_fault() const;
};

Status hip_check(hipError_t e, const char* what);

}  // namespace ptx
