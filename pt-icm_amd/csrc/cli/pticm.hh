// pticm.hh -- C++ facade with the reference's names (namespace pticm) over
// the engine's C ABI.  A reference caller keeps its types and calls:
//   Parameters      include/mpi_defs.hh:12-29
//   Problem         include/problem.hh:7-98 (host model; sweeps run on the GPU)
//   Replica         include/replica.hh:6-66 (its sweep runs on the GPU)
//   PTWorker        include/mpi_defs.hh:33-65 without the MPI fields
//   beta_linspace   src/mpi_utils.cc:6-22
//   get_beta        src/cae_estimator.cc:295-428 (pilot on the GPU, DOS on the host)
//   solve           include/solver.hh:6, src/solver.cc:16-166
//   solve_folder    the folder / restarts CSV mode of src/main.cc:149-208
#pragma once

#include <cstdint>
#include <list>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/ptx.h"

namespace pticm {

typedef std::vector<bool> StateVector;

struct Parameters {
    uint64_t seed = 0;
    uint64_t Nt = 20;
    uint64_t pts = 1000;
    uint64_t pre_samples = 2000;
    uint64_t burn_in = 500;
    uint64_t sweeps = 1;
    double T0 = 0.6;
    double T1 = 2.8;
    double cutoff = 0.0;
    double pre_step = 0.1;
    bool inverse = false;
    bool preprocess = false;
    bool log_data = false;
    bool icm = false;
    uint32_t verbosity = 0;
    // additions of this engine
    uint64_t pre_burn_in = 5000;  // BURN_IN of src/cae_estimator.cc:74
    int32_t ladder_mode = PTX_LADDER_EXACT;
    std::string log_path = "results.csv";  // --log: per-temperature T,C,X,u,m (src/main.cc:493-528)
    std::string ladder_in;                 // read the ladder (one beta per line) instead of making one
    std::string ladder_out;                // write the ladder used
    std::string histograms_out;            // --preprocess: write the pilot's energy histograms
    std::string checkpoint;                // write the engine state after the run
    std::string resume;                    // continue from a state written by --checkpoint
};

// Thrown by the facade when the C ABI reports an error (the reference exits).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

class Problem {
  public:
    Problem() = default;
    explicit Problem(const std::string& path, uint64_t seed = 0, uint64_t burn_in = UINT64_MAX);
    // dense constructor of include/problem.hh:23-39: formulation is N x N
    // row-major, the couplings are its upper triangle (problem.cc:134-138)
    Problem(double* formulation, double* linear, uint64_t seed, uint64_t burn_in, uint64_t N);
    // COO constructor of include/problem.hh:40-62
    Problem(const double* weights, const int* src, const int* dest, const double* linear, uint64_t seed,
            uint64_t burn_in, uint64_t N, uint64_t M);
    double energy(const StateVector& s) const;
    double energy(bool* s) const;
    double maxcut(const StateVector& s) const;
    double calculate_dE(std::vector<double>& dE, const StateVector& s) const;
    // problem.cc:164-181: after accepting a flip of `index` in state s (before
    // the flip is applied), dE[index] changes sign and each neighbour j gets
    // 4 s_index s_j w
    void incremental_dE(std::vector<double>& dE, const StateVector& s, uint32_t index) const;
    uint64_t vcount() const { return info_.n; }
    uint64_t ecount() const { return (uint64_t)info_.m; }
    double get_char_scale() const { return info_.char_scale; }
    bool is_sparse() const { return info_.sparse != 0; }
    uint64_t get_burn_in() const { return burn_in_; }
    uint64_t get_seed() const { return seed_; }
    // the host random stream of include/problem.hh:16-17,70,78 (mt19937_64
    // seeded with the seed, uniform on [0, 1)); the GPU sweeps draw from
    // their own counter-based streams
    double rand_double() { return unif_(rng_); }
    std::mt19937_64& get_rng() { return rng_; }
    const ptx_problem* handle() const { return h_.get(); }

  private:
    struct Csr {
        std::vector<int32_t> rowptr, col, w, h;
    };
    std::shared_ptr<ptx_problem> h_;
    std::shared_ptr<const Csr> csr_;
    ptx_problem_info info_{};
    uint64_t seed_ = 0, burn_in_ = UINT64_MAX;
    std::mt19937_64 rng_;
    std::uniform_real_distribution<double> unif_{0.0, 1.0};
    void init_info();
};

// One replica (include/replica.hh:6-66).  sweep() runs one Metropolis sweep
// of this replica on the GPU (a one-replica engine made on the first sweep
// and again after set_beta or a state written by the caller) and brings
// state, energy, dE, flips and the best state back.  Copies share no device
// state.  A caller that drives many replicas gets the same trajectories,
// far faster, from solve() or the engine's C ABI, which keep every replica
// resident on the device.
class Replica {
  public:
    Replica() = default;
    Replica(Problem& p, double beta, const StateVector& init_state, const std::vector<double>& init_dE,
            double energy, bool log_ene = false, size_t burn_in = 0);
    Replica(const Replica& o) { *this = o; }
    Replica& operator=(const Replica& o);
    void sweep(Problem& p);
    double get_energy() const { return energy; }
    void set_energy(double e) { energy = e; dirty_ = true; }
    void set_ene(double e) { set_energy(e); }
    const StateVector& get_state() const { return state; }
    StateVector& get_state() { dirty_ = true; return state; }  // the caller may write through it
    void set_state(const StateVector& s) { state = s; dirty_ = true; }
    uint64_t get_flips() const { return flips; }
    double get_beta() const { return beta; }
    void set_beta(double b) { beta = b; }
    std::vector<double>& get_dE() { return dE; }
    void set_dE(std::vector<double> d) { dE = std::move(d); }
    const StateVector& get_best_state() const { return best_state; }
    size_t get_best_sweep() const { return best_sweep; }
    double get_best_ene() const { return best_ene; }
    // one entry per sweep after burn_in (the reference logs after every spin
    // attempt, src/replica.cc:10-12)
    const std::list<double>& get_ene_log() const { return ene_log; }

    double energy = 0, beta = 0, best_ene = 1.79769313486231570815e+308;
    uint64_t flips = 0;
    size_t best_sweep = 0;
    StateVector state, best_state;
    std::vector<double> dE;

  private:
    struct Device;
    std::shared_ptr<Device> dev_;  // never shared between copies (operator= drops it)
    bool dirty_ = true;
    bool log_energy_ = false;
    size_t burn_in_ = 0, sweep_counter_ = 0;
    uint64_t serial_ = 0;          // distinguishes the device streams of replicas of one problem
    std::list<double> ene_log;
};

struct PTWorker {
    int wid = 0;
    int numw = 1;
    Parameters params;
    Problem p;
    std::vector<double> beta_values;       // 2 Nt slot betas
    uint64_t swaps = 0;
    uint32_t Nc = 0;
    double characteristic_scale = 0;
    std::vector<uint32_t> beta_indices;    // slot -> replica
    std::vector<uint32_t> swap_attempts;
    std::vector<uint32_t> swap_successes;
    std::vector<Replica> replicas;
    double elapsed = 0;
};

void beta_linspace(double* temps, double T0, double T1, uint64_t num_temp, bool inverse);
// Runs the pilot and the ladder search; fills w.beta_values[0..Nt') and sets
// w.params.Nt = Nt' (src/cae_estimator.cc:425).
void get_beta(PTWorker& w, double estep, double step_size, double r, size_t num_samples);
void solve(PTWorker& w, std::string& problem_path, double& best_ene, size_t& flips, StateVector& best_state);
// The folder / restarts mode of src/main.cc:149-208 as one batched engine per
// restart over every problem file of `folder` (in name order): restart i
// uses seed + i, each problem its own random streams; one CSV row per problem
// and restart with the columns of main.cc:152.
void solve_folder(PTWorker& w, const std::string& folder, const std::string& csv, size_t restarts);
// Prints the result keys of src/solver.cc:151-163.
void print_results(const PTWorker& w, double best_ene, size_t best_sweep, size_t flips, const StateVector& best_state);

}  // namespace pticm
