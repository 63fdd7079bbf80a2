/*
 * oracle.h -- CPU restatement of the PT-ICM hot path (test infrastructure only).
 *
 * This library is the checker for the HIP engine.  Only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * never links or calls it.  It follows the reference line by line where the
 * reference defines the behaviour, and the engine's counter-based random
 * stream where the reference used mt19937_64 (that stream is part of this
 * build's contract, so the oracle defines it identically to the device code).
 *
 * Reference files followed (paths relative to the reference checkout):
 *   src/problem.cc        reader rules, energy, dE, maxcut, incremental dE
 *   src/replica.cc        Metropolis accept rule and best-state tracking
 *   src/replica_exchange.cc  (NN)_a exchange order, decision, counters
 *   src/icm.cc            ICM global-flip rule, seed pick, cluster growth
 *   src/initialization.cc Nc rule, ladder selection, init order
 *   src/mpi_utils.cc:6-22 beta_linspace
 *   src/cae_estimator.cc  pilot run, histograms, multihistogram DOS, ladder
 *   src/solver.cc:67-164  step order and outputs
 *
 * Which functions are pinned by which check is listed in the header of
 * oracle.c next to each function.
 */
#ifndef PTICM_ORACLE_H
#define PTICM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based random numbers and portable FP64 ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_exp(double x);  /* Cody-Waite reduction + Taylor; mul/add only */
double or_log(double x);  /* 2^(-j/32) table reduction + Taylor log1p; mul/add only */
double or_uniform_log(uint32_t r); /* log(r * 2^-32), -inf for r == 0 */

/* ---- problem ---- */
typedef struct or_graph or_graph;
/* Reads a gset / .ising file following src/problem.cc:183-247.  Returns NULL
 * and fills err on failure. */
or_graph* or_read_gset(const char* path, char* err, int errlen);
/* Builds a graph from an undirected COO list (sparse rules of problem.hh:40-62
 * when dense == 0, dense overwrite rules of problem.cc:210-215 otherwise). */
or_graph* or_graph_from_coo(int32_t n, int64_t m, const int32_t* u, const int32_t* v,
                            const double* w, const double* h, int dense, char* err, int errlen);
void or_graph_free(or_graph* g);
int32_t or_graph_n(const or_graph* g);
int64_t or_graph_nnz(const or_graph* g);        /* directed CSR entries */
int64_t or_graph_m_header(const or_graph* g);   /* M as read from the header */
int or_graph_sparse(const or_graph* g);
double or_graph_char_scale(const or_graph* g);
void or_graph_csr(const or_graph* g, int32_t* rowptr, int32_t* col, int32_t* w, int32_t* h);

int64_t or_energy(const or_graph* g, const uint8_t* s);
int64_t or_maxcut(const or_graph* g, const uint8_t* s);
void or_delta_e(const or_graph* g, const uint8_t* s, int64_t* dE);
void or_incremental_de(const or_graph* g, int64_t* dE, const uint8_t* s, int32_t index);

/* Greedy colouring in index order; returns the number of classes. */
int32_t or_colour(const or_graph* g, int32_t* colour);
/* Position order: sorted by (class, index).  class_start has ncls+1 entries. */
void or_positions(const or_graph* g, const int32_t* colour, int32_t ncls,
                  int32_t* pos2orig, int32_t* class_start);

/* ---- ladders ---- */
void or_beta_linspace(double* out, double T0, double T1, int64_t n, int inverse);
int32_t or_icm_count(const double* beta, int32_t nt, double char_scale);
void or_threshold_table(const double* beta, int32_t nt, int32_t fmax, uint32_t* thr);
int32_t or_fmax(const or_graph* g);

/* ---- engine ---- */
enum {
    OR_FLAG_ICM = 1u << 0,
    OR_FLAG_REFERENCE_ORDER = 1u << 1, /* index-order sweep, sequential draws */
    OR_FLAG_PILOT = 1u << 2,           /* one chain, all pairs every step */
    OR_FLAG_NO_EXCHANGE = 1u << 3,
    OR_FLAG_LOG_ENERGY = 1u << 4
};
typedef struct or_engine or_engine;
or_engine* or_engine_create(const or_graph* g, const double* betas, int32_t nt, int32_t chains,
                            uint64_t seed, uint32_t instance, uint32_t flags, int32_t sweeps,
                            int32_t nc_override /* <0: reference Nc rule */);
void or_engine_free(or_engine* e);
void or_engine_step(or_engine* e, int64_t nsteps);
int32_t or_engine_nrep(const or_engine* e);
int32_t or_engine_nc(const or_engine* e);
void or_engine_spins(const or_engine* e, uint8_t* out);        /* [nrep][n] index order */
void or_engine_set_spins(or_engine* e, const uint8_t* in);     /* recomputes energies */
void or_engine_energies(const or_engine* e, int64_t* out);     /* [nrep] */
void or_engine_slot2rep(const or_engine* e, int32_t* out);     /* [nrep] */
void or_engine_swap_stats(const or_engine* e, uint32_t* att, uint32_t* succ, uint64_t* swaps);
void or_engine_flips(const or_engine* e, uint64_t* out);       /* [nrep] */
void or_engine_best(const or_engine* e, int64_t* bestE, int64_t* best_sweep, uint8_t* best_state);
/* Energy log: entry [s][slot] = energy of the replica at slot after the sweeps
 * of step s (src/cae_estimator.cc:117-121 logs at the same point). */
int64_t or_engine_log_len(const or_engine* e);
void or_engine_energy_log(const or_engine* e, int32_t* out);
/* ICM trace of the last step: per target t (nt entries): seed position or -1,
 * cluster size, global-flip flag. */
void or_engine_icm_trace(const or_engine* e, int32_t* seed, int32_t* size, int32_t* gflip);

/* ---- multihistogram DOS and constant-acceptance ladder (FP64) ---- */
/* make_histogram (src/cae_estimator.cc:45-69) of one energy log with estep. */
int64_t or_make_histogram(const int32_t* elog, int64_t n, int64_t stride, double estep,
                          double* emin, double* emax, uint32_t* hist, int64_t cap);
/* get_beta from per-slot energy logs (src/cae_estimator.cc:295-428, with the
 * pilot done elsewhere).  elog is [nsamples][nt].  Returns the number of
 * temperatures selected; betas_out has room for nt entries. */
int32_t or_get_beta_from_logs(const int32_t* elog, int64_t nsamples, int32_t nt,
                              const double* hist_betas, double estep, double step_size,
                              double r, double beta_start, double beta_end,
                              double* betas_out, double* logdos_out, int64_t logdos_cap,
                              int64_t* nbins_out, double* logz_out);
double or_acceptance_rate(const double* logdos, const double* p1, const double* p2, int64_t nb,
                          double beta_1, double beta_2, double estep, double emin);
double or_partition_func(const double* logdos, int64_t nb, double beta, double emin, double estep);
/* Pilot run following gather_samples (src/cae_estimator.cc:76-151) on the
 * engine stream: one chain, burn_in + nsamples steps; writes the logged
 * energies [nsamples][nt]. */
void or_gather_samples(const or_graph* g, const double* hist_betas, int32_t nt, uint64_t seed,
                       uint32_t instance, int64_t burn_in, int64_t nsamples, int32_t* elog,
                       uint32_t extra_flags);

#ifdef __cplusplus
}
#endif
#endif
