// ladder.hpp -- host FP64 temperature ladders: linear ladders, the ICM slot
// rule, the Metropolis threshold tables, and the multihistogram (Ferrenberg-
// Swendsen) density of states with the constant-acceptance ladder search of
// src/cae_estimator.cc.
#pragma once

#include <cstdint>
#include <vector>

namespace ptx {

// src/mpi_utils.cc:6-22
void beta_linspace(double* out, double T0, double T1, int64_t n, bool inverse);
// src/initialization.cc:158-164
int32_t icm_count(const double* beta, int32_t nt, double char_scale);
// hist_betas of src/cae_estimator.cc:309-328
void pilot_betas(double T0, double T1, int32_t nt, bool inverse, double* out, double* beta_start, double* beta_end);
// thr[t][m-1] = floor(exp(-beta_t * 2m) * 2^32), saturated at 2^32-1, with
// the portable exp; accept an uphill dE = 2m iff the 32-bit draw is below.
void threshold_table(const double* beta, int32_t nt, int32_t fmax, uint32_t* thr);

// Per-slot histograms in the reference layout (src/cae_estimator.cc:45-69).
struct SlotHist {
    double emin = 0, emax = 0;
    std::vector<uint32_t> h;
};
// Converts the engine layout (energy = e_base + 2*bin) into make_histogram's.
std::vector<SlotHist> rebin(const uint32_t* counts, int32_t nt, int64_t nbins, int64_t e_base, double estep);

struct Dos {
    std::vector<double> logdos;  // bins of width estep from min_emin
    std::vector<double> logz;    // per histogram temperature
    double min_emin = 0, max_emax = 0, estep = 2;
};
// get_partition_iterate + calculate_dos (src/cae_estimator.cc:155-253)
Dos solve_dos(const std::vector<SlotHist>& hist, const double* hist_betas, double estep);
// The same equations solved by damped Newton (converges in a few iterations
// where the reference's fixed-point iteration needs thousands).
Dos solve_dos_newton(const std::vector<SlotHist>& hist, const double* hist_betas, double estep);

double partition_func(const std::vector<double>& logdos, double beta, double emin, double estep);
// acceptance_rate (src/cae_estimator.cc:262-293); fast = O(nbins) prefix form.
double acceptance_rate(const std::vector<double>& p1, const std::vector<double>& p2, double beta_1, double beta_2,
                       double estep, double emin, bool fast);
// Ladder search (src/cae_estimator.cc:360-425).  Returns the count chosen.
int32_t ladder_search(const Dos& dos, int32_t nt, double beta_start, double beta_end, double step_size, double r,
                      bool fast, double* betas_out);

}  // namespace ptx
