// pticm.hh -- C++ facade with the reference's names (namespace pticm) over
// the engine's C ABI.  A reference caller keeps its types and calls:
//   Parameters      include/mpi_defs.hh:12-29
//   Problem         include/problem.hh:7-98 (host model; sweeps run on the GPU)
//   Replica         include/replica.hh:6-66 (read-only snapshot of a replica)
//   PTWorker        include/mpi_defs.hh:33-65 without the MPI fields
//   beta_linspace   src/mpi_utils.cc:6-22
//   get_beta        src/cae_estimator.cc:295-428 (pilot on the GPU, DOS on the host)
//   solve           include/solver.hh:6, src/solver.cc:16-166
#pragma once

#include <cstdint>
#include <memory>
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
    // COO constructor of include/problem.hh:40-62
    Problem(const double* weights, const int* src, const int* dest, const double* linear, uint64_t seed,
            uint64_t burn_in, uint64_t N, uint64_t M);
    double energy(const StateVector& s) const;
    double maxcut(const StateVector& s) const;
    double calculate_dE(std::vector<double>& dE, const StateVector& s) const;
    uint64_t vcount() const { return info_.n; }
    uint64_t ecount() const { return (uint64_t)info_.m; }
    double get_char_scale() const { return info_.char_scale; }
    bool is_sparse() const { return info_.sparse != 0; }
    uint64_t get_burn_in() const { return burn_in_; }
    const ptx_problem* handle() const { return h_.get(); }

  private:
    std::shared_ptr<ptx_problem> h_;
    ptx_problem_info info_{};
    uint64_t seed_ = 0, burn_in_ = UINT64_MAX;
    void init_info();
};

class Replica {
  public:
    double get_energy() const { return energy; }
    double get_beta() const { return beta; }
    uint64_t get_flips() const { return flips; }
    const StateVector& get_state() const { return state; }
    const StateVector& get_best_state() const { return best_state; }
    size_t get_best_sweep() const { return best_sweep; }
    double get_best_ene() const { return best_ene; }
    double energy = 0, beta = 0, best_ene = 0;
    uint64_t flips = 0;
    size_t best_sweep = 0;
    StateVector state, best_state;
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
// Prints the result keys of src/solver.cc:151-163.
void print_results(const PTWorker& w, double best_ene, size_t best_sweep, size_t flips, const StateVector& best_state);

}  // namespace pticm
