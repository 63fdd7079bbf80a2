/*
 * ptx.h -- C ABI of the MI355X parallel-tempering + ICM engine (libptx.so).
 *
 * Plain pointers and sizes only; opaque handles; every entry point returns a
 * ptx_status (0 = success, negative = error, never exits the process); the
 * message of the last error on the calling thread is available through
 * ptx_last_error.  Buffers are caller-owned host memory.  An engine owns one
 * HIP stream and is not thread-safe; distinct engines are independent.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * mxburns2022/PT-ICM).
 */
#ifndef PTX_H
#define PTX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ptx_status;
#define PTX_OK 0
#define PTX_ERR_INVALID (-1)     /* bad argument or malformed input */
#define PTX_ERR_IO (-2)          /* file could not be read */
#define PTX_ERR_UNSUPPORTED (-3) /* input outside the engine's limits */
#define PTX_ERR_DEVICE (-4)      /* HIP runtime error (no device, launch failure) */
#define PTX_ERR_NOMEM (-5)       /* host or device allocation failed */

typedef struct ptx_problem ptx_problem;
typedef struct ptx_engine ptx_engine;

/* Copies the last error message of this thread into buf (NUL-terminated). */
ptx_status ptx_last_error(char* buf, size_t len);
/* Library version string and the device it would run on ("" if none). */
const char* ptx_version(void);
int32_t ptx_device_count(void);

/* ------------------------------------------------------------------ */
/* Problem (replaces pticm::Problem, include/problem.hh:7-98)          */
/* ------------------------------------------------------------------ */

/* Problem(std::string path, seed, burn_in), src/problem.cc:183-247: gset or
 * .ising text; 0-based iff the path contains ".gset" or ".ising", else
 * 1-based; missing weight = 1; sparse iff M / (N(N-1)/2) < 0.1.  Weights and
 * linear terms must be integers (PTX_ERR_UNSUPPORTED otherwise). */
ptx_status ptx_problem_read_gset(const char* path, ptx_problem** out);

/* Problem(double* weights, int* src, int* dest, double* linear, seed, burn_in,
 * N, M), include/problem.hh:40-62 (dense == 0), and the dense-matrix rules of
 * src/problem.cc:210-215 (dense != 0: last weight of a pair wins, zero = no
 * edge).  h may be NULL (all zero). */
ptx_status ptx_problem_create_coo(int32_t n, int64_t m, const int32_t* src, const int32_t* dst,
                                  const double* w, const double* h, int32_t dense, ptx_problem** out);
void ptx_problem_destroy(ptx_problem* p);

typedef struct {
    int32_t n;            /* vcount(), problem.hh:67 */
    int64_t m;            /* ecount(), problem.hh:68 (header M) */
    int64_t nnz;          /* directed adjacency entries */
    int32_t sparse;       /* is_sparse(), problem.hh:74 */
    double char_scale;    /* get_char_scale(), problem.hh:66 */
    int32_t n_colours;    /* colour classes (DSatur, then sizes balanced towards ceil(n / n_colours)) */
    int32_t fmax;         /* max_i |h_i| + sum_j |w_ij| */
    int64_t e_min_bound;  /* -(sum |w| over edges + sum |h|) */
    int64_t e_max_bound;
    int32_t has_field;    /* any h_i != 0 */
    int32_t lattice_dim;  /* 2 or 3 for a periodic L^d lattice in index order (rows exactly the
                             wrapped axis neighbours): cluster moves grow bit-set clusters; else 0 */
    int32_t lattice_L;
} ptx_problem_info;
ptx_status ptx_problem_info_get(const ptx_problem* p, ptx_problem_info* out);

/* Problem::energy, src/problem.cc:121-142.  state: n bytes, 0/1. */
ptx_status ptx_problem_energy(const ptx_problem* p, const uint8_t* state, int64_t* energy);
/* Problem::maxcut, src/problem.cc:145-162. */
ptx_status ptx_problem_maxcut(const ptx_problem* p, const uint8_t* state, int64_t* cut);
/* Problem::calculate_dE, src/problem.cc:33-63: dE[n] and the energy. */
ptx_status ptx_problem_delta_e(const ptx_problem* p, const uint8_t* state, int64_t* dE, int64_t* energy);
/* The colour-class order the engine sweeps in: pos2orig[n], class_start[n_colours+1]. */
ptx_status ptx_problem_colouring(const ptx_problem* p, int32_t* pos2orig, int32_t* class_start);
/* Adjacency in index order: rowptr[n+1], col[nnz], w[nnz], h[n]. */
ptx_status ptx_problem_csr(const ptx_problem* p, int32_t* rowptr, int32_t* col, int32_t* w, int32_t* h);

/* ------------------------------------------------------------------ */
/* Ladders (host, FP64)                                                */
/* ------------------------------------------------------------------ */

/* beta_linspace, src/mpi_utils.cc:6-22 (linear in beta, ascending). */
ptx_status ptx_beta_linspace(double* out, double T0, double T1, int64_t n, int32_t inverse);
/* Nc rule of initialize_temperatures, src/initialization.cc:158-164. */
int32_t ptx_icm_count(const double* beta, int32_t nt, double char_scale);
/* hist_betas of get_beta, src/cae_estimator.cc:309-328: nt values from
 * max(b0,b1) down to min(b0,b1) by accumulated steps; returns both ends. */
ptx_status ptx_pilot_betas(double T0, double T1, int32_t nt, int32_t inverse, double* out,
                           double* beta_start, double* beta_end);

#define PTX_LADDER_EXACT 0 /* the reference's O(nbins^2) acceptance sum */
#define PTX_LADDER_FAST 1  /* O(nbins) prefix-sum form of the same sum */
#define PTX_LADDER_NEWTON 2 /* FAST, with the DOS equations solved by Newton instead of fixed-point iteration */
/* get_beta after the pilot, src/cae_estimator.cc:330-428: multihistogram DOS
 * (Ferrenberg-Swendsen in log space) from per-slot energy histograms, then the
 * constant-acceptance ladder search toward rate r.  counts is [nt][nbins] in
 * the engine's histogram layout (energy = e_base + 2*bin); estep as the
 * reference (2).  betas_out holds nt entries; *nt_out = temperatures chosen. */
ptx_status ptx_ladder_from_histograms(const uint32_t* counts, int32_t nt, int64_t nbins, int64_t e_base,
                                      const double* hist_betas, double estep, double step_size, double r,
                                      double beta_start, double beta_end, int32_t mode, double* betas_out,
                                      int32_t* nt_out);
/* Predicted swap acceptance between adjacent ladder entries from the same
 * DOS (used to judge a proposed ladder against target +-0.05). */
ptx_status ptx_ladder_predict_rates(const uint32_t* counts, int32_t nt, int64_t nbins, int64_t e_base,
                                    const double* hist_betas, double estep, const double* ladder,
                                    int32_t n_ladder, int32_t mode, double* rates_out);

/* ------------------------------------------------------------------ */
/* Engine (replaces the solve() loop, src/solver.cc:16-166, with       */
/* Replica::sweep replica.cc:5-38, try_replica_exchange                */
/* replica_exchange.cc:5-100, icm icm.cc:8-215, gather_samples         */
/* cae_estimator.cc:76-151)                                            */
/* ------------------------------------------------------------------ */

#define PTX_FLAG_ICM (1u << 0)         /* isoenergetic cluster moves (--icm) */
#define PTX_FLAG_PILOT (1u << 2)       /* pilot run: one chain, all pairs per step */
#define PTX_FLAG_NO_EXCHANGE (1u << 3) /* sweeps only (testing) */
#define PTX_FLAG_HISTOGRAM (1u << 5)   /* accumulate per-slot energy histograms */
#define PTX_FLAG_OBSERVE (1u << 6)     /* accumulate per-slot observables (--log); split mode only */

#define PTX_MODE_AUTO 0
#define PTX_MODE_SPLIT 1 /* sweep / exchange / ICM kernels per step */
#define PTX_MODE_FUSED 2 /* one workgroup per instance runs whole PT steps */
#define PTX_MODE_COOP 3  /* one workgroup per replica of every instance, all resident, one wait per step */

#define PTX_CSR_AUTO 0   /* adjacency in LDS when it fits, else read through L2 */
#define PTX_CSR_LDS 1
#define PTX_CSR_GLOBAL 2

typedef struct {
    int32_t chains;          /* 2 for the main run (initialization.cc:166), 1 for a pilot */
    int32_t sweeps;          /* --sweeps */
    uint32_t flags;          /* PTX_FLAG_* */
    int32_t nc;              /* ICM slots; -1 = reference rule per instance */
    int64_t hist_start;      /* first step whose energies enter the histogram */
    int32_t mode;            /* PTX_MODE_* */
    int32_t waves_per_group; /* split: replicas (one wave each) per workgroup, 1/2/4/8; fused: waves per
                                instance, 4/8/16; cooperative: waves per replica, 4, 8 or 16 (16 with one
                                position per lane and no ICM slot, else 8) (0 = auto) */
    int32_t steps_per_launch;/* fused mode: PT steps per launch (0 = auto) */
    int32_t use_graph;       /* fused mode: G > 0 captures G launches of steps_per_launch steps in a
                                hipGraph and replays it (0 = eager launches) */
    int32_t spins_per_lane;  /* cooperative mode: 1 or 4 positions per lane (0 = auto) */
    int32_t csr_placement;   /* fused mode: PTX_CSR_AUTO, PTX_CSR_LDS or PTX_CSR_GLOBAL */
    int32_t icm_waves;       /* split mode: waves per ICM slot pair, 1, 2 or 4 (0 = 1) */
    int32_t small_batch_coop;/* split mode: while every replica of the active instances fits one
                                cooperative launch (for example the last instances of a batch whose
                                others were retired), run them with the cooperative kernel */
    int32_t coop_wait_us;    /* cooperative launches: bound of one per-step wait for the other replicas
                                of an instance, in microseconds (0 = 100,000).  A wait that exceeds it
                                ends the launch; the next ptx_engine_step, ptx_engine_sync or read then
                                returns PTX_ERR_DEVICE and no further launch is enqueued.  A negative
                                value gives up at the first poll that finds a replica missing (tests of
                                that error path). */
    int32_t sliced_sweep;    /* split mode: 0 (auto) sweeps the instances with couplings +-1, no field and
                                degree <= 63 that take no multispin lattice sweep 32 replicas at a time
                                (a workgroup per 32 replicas, each neighbour word one gather for all 32,
                                counts bit-sliced; the same decisions) with 4 waves per workgroup, 1, 2,
                                4 or 8 with that many; -1: every instance by the per-replica sweep */
} ptx_engine_config;
void ptx_engine_config_default(ptx_engine_config* cfg);

/* Creates an engine over n_instances independent problems.  Instance i has
 * n_temps[i] temperatures per chain, its ladder in betas starting at
 * sum_{j<i} n_temps[j].  All replicas start from Bernoulli(1/2) states drawn
 * from the counter-based stream (seed, instance). */
ptx_status ptx_engine_create(const ptx_problem* const* problems, int32_t n_instances, const double* betas,
                             const int32_t* n_temps, uint64_t seed, const ptx_engine_config* cfg,
                             ptx_engine** out);
void ptx_engine_destroy(ptx_engine* e);

/* Launch-shape autotuner (no counterpart in the reference, which has one
 * CPU code path).  Times `steps` PT steps (after a warm-up launch) of every
 * launch shape that fits the batch -- execution mode, waves (replicas) per
 * workgroup, positions per lane, adjacency in LDS or read through L2, waves
 * per cluster move -- and writes the fastest into *best (a copy of *base
 * with mode, waves_per_group, spins_per_lane, csr_placement and icm_waves
 * set).  The random streams are indexed by position, replica and step, so
 * every shape yields the same trajectories.
 * trials (may be NULL) receives up to max_trials measured shapes, *n_trials
 * their count. */
typedef struct {
    int32_t mode, waves_per_group, spins_per_lane, csr_placement, icm_waves;
    double us_per_step;
} ptx_tune_trial;
ptx_status ptx_engine_autotune(const ptx_problem* const* problems, int32_t n_instances, const double* betas,
                               const int32_t* n_temps, uint64_t seed, const ptx_engine_config* base, int32_t steps,
                               ptx_engine_config* best, ptx_tune_trial* trials, int32_t max_trials,
                               int32_t* n_trials);
/* Enqueues n_steps PT steps (asynchronous; errors of the launch are reported). */
ptx_status ptx_engine_step(ptx_engine* e, int64_t n_steps);
/* Retires (active[i] == 0) or reactivates instances: later steps run only the
 * active ones, with their work lists rebuilt, and retired instances keep
 * their state.  A batch whose instances finish at different times (for
 * example on reaching a target energy) stops paying for the finished ones.
 * No counterpart in the reference (one instance per run).  Every instance
 * keeps its own PT step counter (the index of its random streams), so a
 * reactivated instance continues exactly where it stopped. */
ptx_status ptx_engine_set_active(ptx_engine* e, const uint8_t* active);
ptx_status ptx_engine_sync(ptx_engine* e);

typedef struct {
    int32_t n_instances;
    int64_t steps_done;
    int32_t mode;            /* resolved mode (after steps: the mode of the latest launches) */
    int32_t waves_per_group; /* resolved */
    int32_t steps_per_launch;
    int64_t total_replicas;
    int64_t spin_updates_per_step; /* sum over replicas of n * sweeps */
    int64_t lds_bytes_per_step;    /* algorithmic LDS bytes of the sweep (4 d_i + 6 per update) */
    int32_t csr_in_lds;            /* fused and cooperative modes: 1 if the adjacency is staged in LDS
                                      (cooperative: in the latest launch) */
    int32_t icm_waves;             /* waves that grow one cluster (split: per pair; fused and
                                      cooperative: 1) */
    int64_t coop_capacity;         /* workgroups (replicas) one cooperative launch can hold; 0: none */
} ptx_engine_info;
ptx_status ptx_engine_info_get(const ptx_engine* e, ptx_engine_info* out);

typedef struct {
    int32_t n, nt, nrep, nc, n_colours;
    int32_t hist_bins;
    int64_t e_base;
    int32_t lattice_sweep;   /* 2 or 3: split-mode sweeps of this instance use the multispin lattice
                                sweep (periodic lattice, even L, couplings +-1, no field); 0: gathers */
    int32_t sliced_sweep;    /* 1: split-mode sweeps of this instance are replica-sliced (32 replicas per
                                wave, see ptx_engine_config.sliced_sweep); 0: one replica per wave */
} ptx_instance_info;
ptx_status ptx_engine_instance_info(const ptx_engine* e, int32_t inst, ptx_instance_info* out);

/* Readers (synchronize the engine stream first).  Replica order; spins in the
 * problem's index order, one byte per spin. */
ptx_status ptx_engine_read_spins(ptx_engine* e, int32_t inst, uint8_t* out);        /* [nrep][n] */
ptx_status ptx_engine_write_spins(ptx_engine* e, int32_t inst, const uint8_t* in);  /* recomputes energies */
ptx_status ptx_engine_read_energies(ptx_engine* e, int32_t inst, int64_t* out);     /* [nrep] */
ptx_status ptx_engine_read_perm(ptx_engine* e, int32_t inst, int32_t* slot2rep);    /* [nrep] */
ptx_status ptx_engine_read_swap_stats(ptx_engine* e, int32_t inst, uint32_t* attempts, uint32_t* successes,
                                      uint64_t* swaps);
ptx_status ptx_engine_read_flips(ptx_engine* e, int32_t inst, uint64_t* flips);     /* [nrep] */
/* Per-replica best energy, the sweep it was reached in, and the state
 * (replica.cc:23-27; checked after every colour class that accepted a flip). */
ptx_status ptx_engine_read_best(ptx_engine* e, int32_t inst, int64_t* best_e, int64_t* best_sweep,
                                uint8_t* best_state);
/* Histogram counts [nrep_slots][hist_bins], energy = e_base + 2*bin. */
ptx_status ptx_engine_read_histograms(ptx_engine* e, int32_t inst, uint32_t* counts);
/* Per-slot observables of --log (the writer of src/main.cc:490-530, whose
 * samples were logged by Replica::sweep, src/replica.cc:10-12), accumulated
 * by an engine created with PTX_FLAG_OBSERVE after every PT step from step
 * hist_start on, on the state at the end of the step.  up = number of spins
 * set (the reference's magnetisation m is up / n: 0.5 means no net moment);
 * sums are exact (integers, the squares in FP64 below 2^53).  From them:
 * u = mean_e / n, C = beta^2 / n (mean_e2 - mean_e^2), chi = beta n var(up / n),
 * m = mean_up / n. */
typedef struct {
    int64_t samples;
    int64_t sum_e;       /* sum of E */
    double sum_e2;       /* sum of E^2 */
    int64_t sum_up;      /* sum of the number of spins set */
    double sum_up2;      /* sum of its square */
    int64_t sum_abs_m;   /* sum of |2 up - n| (absolute magnetisation times n) */
} ptx_slot_observables;
ptx_status ptx_engine_read_observables(ptx_engine* e, int32_t inst, ptx_slot_observables* out); /* [nrep] */
/* Lowest best energy of every instance over its replicas, out[n_instances]
 * (INT64_MAX where nothing is recorded yet), in one call: the poll of
 * time-to-target loops.  Best energies as in ptx_engine_read_best. */
ptx_status ptx_engine_read_min_energies(ptx_engine* e, int64_t* out_per_instance);

/* Checkpoint and resume (no counterpart in the reference, whose runs cannot
 * be stopped and resumed).  The state is every array a PT step changes --
 * spins, energies, best energies, sweeps and states, the temperature
 * permutation, flip and swap counters, histograms, observables -- plus each
 * instance's step counter and active flag, in a caller-owned buffer of
 * ptx_engine_state_size bytes (host memory; write it to a file as is).  It
 * loads into an engine created with the same problems, ladders, flags,
 * sweeps, hist_start and seed (checked: PTX_ERR_INVALID otherwise), in any
 * execution mode, and the loaded engine continues bit for bit as the saved
 * one would have: the random streams are indexed by (seed, instance, step). */
ptx_status ptx_engine_state_size(const ptx_engine* e, int64_t* bytes);
ptx_status ptx_engine_save_state(ptx_engine* e, void* buf, int64_t bytes);
ptx_status ptx_engine_load_state(ptx_engine* e, const void* buf, int64_t bytes);

/* Device time of the kernels enqueued since the last reset, measured with
 * HIP events on the engine stream: total ms and launch count per kernel. */
typedef struct {
    double sweep_ms, exchange_ms, icm_ms, fused_ms;
    int64_t sweep_launches, exchange_launches, icm_launches, fused_launches;
} ptx_engine_timing;
ptx_status ptx_engine_timing_enable(ptx_engine* e, int32_t on);
ptx_status ptx_engine_timing_get(ptx_engine* e, ptx_engine_timing* out);

/* Fused and cooperative modes: wall-clock microseconds that workgroup 0
 * spent in the sweep (us[0]), exchange (us[1]) and cluster-move (us[2])
 * phases, and (cooperative mode) publishing states and waiting at the grid
 * barrier (us[3]), since profiling was switched on (on >= 0 resets and
 * switches; on < 0 only reads).  us may be NULL, else holds 4 values. */
ptx_status ptx_engine_phase_profile(ptx_engine* e, int32_t on, double* us);

/* Device self-test of the FP64 helpers the exchange uses (exp/log portable
 * forms), for bitwise comparison against the host forms. */
ptx_status ptx_selftest_fp64(const double* x, int64_t n, double* exp_out, double* log_out);
ptx_status ptx_host_fp64(const double* x, int64_t n, double* exp_out, double* log_out);
/* Device Philox4x32-10 for counters ctr[n][4] under key[2] -> out[n][4]. */
ptx_status ptx_selftest_philox(const uint32_t* ctr, int64_t n, const uint32_t* key, uint32_t* out);
/* Microbenchmarks that pin the roofline peaks: a device copy of nbytes
 * (k_copy16; also calibrates memory-traffic counters) and an LDS read
 * bandwidth kernel (k_lds_read, 16-byte reads, 256 threads per block). */
ptx_status ptx_selftest_copy(int64_t nbytes, int32_t reps, double* ms_per_copy);
ptx_status ptx_selftest_lds_bandwidth(int32_t blocks, int32_t iters, double* gbs);

/* ------------------------------------------------------------------ */
/* Ladder surrogate: a fully connected network (ReLU hidden layers,   */
/* softplus outputs) trained and evaluated with BF16 MFMA kernels and */
/* FP32 accumulation.  No counterpart in the reference: the ladder it */
/* predicts stands in for the one get_beta (src/cae_estimator.cc:     */
/* 295-428) computes from a long pilot run, which remains the         */
/* fallback.  Parameters are flat FP32: per layer W[in][out] (row-    */
/* major) then b[out].  An MLP owns one HIP stream; not thread-safe.  */
/* ------------------------------------------------------------------ */
typedef struct ptx_mlp ptx_mlp;
/* dims[0..n_dims-1]: layer widths (input first); weights He-uniform from seed. */
ptx_status ptx_mlp_create(const int32_t* dims, int32_t n_dims, int32_t max_batch, uint64_t seed, ptx_mlp** out);
void ptx_mlp_destroy(ptx_mlp* m);
ptx_status ptx_mlp_param_count(const ptx_mlp* m, int64_t* count);
ptx_status ptx_mlp_set_params(ptx_mlp* m, const float* params, int64_t count);
ptx_status ptx_mlp_get_params(ptx_mlp* m, float* params, int64_t count);
/* Gradients of the last ptx_mlp_loss_grad (same layout as the parameters). */
ptx_status ptx_mlp_get_grads(ptx_mlp* m, float* grads, int64_t count);
/* y[batch][out] = softplus(network(x[batch][in])). */
ptx_status ptx_mlp_forward(ptx_mlp* m, const float* x, int32_t batch, float* y);
/* ptx_mlp_forward without any allocation, for callers that cannot allocate
 * device memory (serving loops, harnesses that forbid hipMalloc); no
 * counterpart in the reference.  The network is given by its widths and flat
 * parameters (the layout of ptx_mlp_set_params) instead of a handle; x, y as
 * in ptx_mlp_forward, and the result equals ptx_mlp_forward's bit for bit.
 * Works in a fixed 4 MiB device array per GPU (held until the process ends):
 * parameters (FP32 and BF16 copies) plus the batch's activations must fit,
 * else PTX_ERR_UNSUPPORTED (the 516-256-256-256-64 surrogate fits batches up
 * to 288 rows).  Calls on one GPU share the array and are serialised
 * by a lock; each call writes everything it reads.  Synchronous. */
ptx_status ptx_mlp_forward_noalloc(const int32_t* dims, int32_t n_dims, const float* params, int64_t param_count,
                                   const float* x, int32_t batch, float* y);
/* Loss = (1/batch) sum over outputs of: mode 1 (y - t)^2, mode 2
 * max(0, t - y)^2, mode 0 nothing; computes all gradients on the device. */
ptx_status ptx_mlp_loss_grad(ptx_mlp* m, const float* x, const float* target, const uint8_t* mode, int32_t batch,
                             double* loss);
/* One Adam step with the gradients of the last loss_grad. */
ptx_status ptx_mlp_adam(ptx_mlp* m, float lr, float beta1, float beta2, float eps);
/* Minibatch training with the data resident on the device: epochs passes of
 * shuffled batches (a remainder smaller than batch is skipped per epoch);
 * epoch_loss[epochs] receives the mean batch loss of each epoch (may be NULL). */
ptx_status ptx_mlp_train(ptx_mlp* m, const float* x, const float* target, const uint8_t* mode, int64_t n_samples,
                         int32_t batch, int32_t epochs, float lr, uint64_t seed, double* epoch_loss);
/* C[M][N] = A[M][K] . B[K][N] with BF16 operands (raw bits) on the MFMA
 * kernel the network uses; M, N multiples of 32, K of 8. */
ptx_status ptx_selftest_gemm_bf16(int32_t M, int32_t N, int32_t K, const uint16_t* A, const uint16_t* B, float* C);

#ifdef __cplusplus
}
#endif
#endif
