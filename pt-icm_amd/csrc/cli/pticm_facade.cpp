// pticm_facade.cpp -- see pticm.hh.
#include <chrono>
#include <cstdio>
#include <iostream>

#include "pticm.hh"

namespace pticm {

namespace {

void check(ptx_status s) {
    if (s == PTX_OK) return;
    char buf[512];
    ptx_last_error(buf, sizeof buf);
    throw Error(s, buf);
}

struct EngineHandle {
    ptx_engine* e = nullptr;
    ~EngineHandle() { ptx_engine_destroy(e); }
};

std::vector<uint8_t> to_bytes(const StateVector& s) {
    std::vector<uint8_t> b(s.size());
    for (size_t i = 0; i < s.size(); ++i) b[i] = s[i] ? 1 : 0;
    return b;
}

void write_update(const PTWorker& w, const std::string& msg, uint32_t level) {  // utils.hh:71-77
    if (w.wid == 0 && level <= w.params.verbosity) {
        std::printf("%s\n", msg.c_str());
        std::fflush(stdout);
    }
}

std::string fmt_double(double x) {  // fmt's "{}" for the integral energies printed here
    char buf[64];
    if (x == (double)(long long)x) std::snprintf(buf, sizeof buf, "%lld", (long long)x);
    else std::snprintf(buf, sizeof buf, "%.17g", x);
    return buf;
}

}  // namespace

Problem::Problem(const std::string& path, uint64_t seed, uint64_t burn_in) : seed_(seed), burn_in_(burn_in) {
    ptx_problem* p = nullptr;
    check(ptx_problem_read_gset(path.c_str(), &p));
    h_.reset(p, ptx_problem_destroy);
    init_info();
}

Problem::Problem(const double* weights, const int* src, const int* dest, const double* linear, uint64_t seed,
                 uint64_t burn_in, uint64_t N, uint64_t M)
    : seed_(seed), burn_in_(burn_in) {
    ptx_problem* p = nullptr;
    check(ptx_problem_create_coo((int32_t)N, (int64_t)M, src, dest, weights, linear, 0, &p));
    h_.reset(p, ptx_problem_destroy);
    init_info();
}

void Problem::init_info() { check(ptx_problem_info_get(h_.get(), &info_)); }

double Problem::energy(const StateVector& s) const {
    int64_t e = 0;
    auto b = to_bytes(s);
    check(ptx_problem_energy(h_.get(), b.data(), &e));
    return (double)e;
}

double Problem::maxcut(const StateVector& s) const {
    int64_t c = 0;
    auto b = to_bytes(s);
    check(ptx_problem_maxcut(h_.get(), b.data(), &c));
    return (double)c;
}

double Problem::calculate_dE(std::vector<double>& dE, const StateVector& s) const {
    auto b = to_bytes(s);
    std::vector<int64_t> d(s.size());
    int64_t e = 0;
    check(ptx_problem_delta_e(h_.get(), b.data(), d.data(), &e));
    dE.assign(d.begin(), d.end());
    return (double)e;
}

void beta_linspace(double* temps, double T0, double T1, uint64_t num_temp, bool inverse) {
    check(ptx_beta_linspace(temps, T0, T1, (int64_t)num_temp, inverse ? 1 : 0));
}

void get_beta(PTWorker& w, double estep, double step_size, double r, size_t num_samples) {
    const int32_t nt = (int32_t)w.params.Nt;
    std::vector<double> hist_betas(nt);
    double bs = 0, be = 0;
    check(ptx_pilot_betas(w.params.T0, w.params.T1, nt, w.params.inverse ? 1 : 0, hist_betas.data(), &bs, &be));
    write_update(w, "Gathering initial samples...", 1);
    ptx_engine_config cfg;
    ptx_engine_config_default(&cfg);
    cfg.chains = 1;
    cfg.flags = PTX_FLAG_PILOT | PTX_FLAG_HISTOGRAM;
    cfg.hist_start = (int64_t)w.params.pre_burn_in;
    const ptx_problem* probs[1] = {w.p.handle()};
    EngineHandle pilot;
    check(ptx_engine_create(probs, 1, hist_betas.data(), &nt, w.params.seed, &cfg, &pilot.e));
    check(ptx_engine_step(pilot.e, (int64_t)(w.params.pre_burn_in + num_samples)));
    ptx_instance_info ii;
    check(ptx_engine_instance_info(pilot.e, 0, &ii));
    std::vector<uint32_t> counts((size_t)ii.nrep * ii.hist_bins);
    check(ptx_engine_read_histograms(pilot.e, 0, counts.data()));
    write_update(w, "Determining partition functions...", 1);
    std::vector<double> out(nt);
    int32_t chosen = 0;
    check(ptx_ladder_from_histograms(counts.data(), nt, ii.hist_bins, ii.e_base, hist_betas.data(), estep, step_size,
                                     r, bs, be, w.params.ladder_mode, out.data(), &chosen));
    w.beta_values.assign(2 * (size_t)chosen, 0.0);
    for (int32_t i = 0; i < chosen; ++i) w.beta_values[i] = out[i];
    w.params.Nt = (uint64_t)chosen;
    std::cout << chosen << " replica temperatures selected" << std::endl;
}

void solve(PTWorker& w, std::string& problem_path, double& best_ene, size_t& flips, StateVector& best_state) {
    const auto t0 = std::chrono::steady_clock::now();
    w.p = Problem(problem_path, w.params.seed, w.params.burn_in);
    w.characteristic_scale = w.p.get_char_scale();
    write_update(w, "Initializing Temperatures", 1);
    if (w.params.preprocess) {
        get_beta(w, 2, w.params.pre_step, 0.4, w.params.pre_samples);
    } else {
        w.beta_values.assign(2 * w.params.Nt, 0.0);
        beta_linspace(w.beta_values.data(), w.params.T0, w.params.T1, w.params.Nt, w.params.inverse);
    }
    const int32_t nt = (int32_t)w.params.Nt;
    for (int32_t i = 0; i < nt; ++i) w.beta_values[nt + i] = w.beta_values[i];
    w.Nc = (uint32_t)ptx_icm_count(w.beta_values.data(), nt, w.characteristic_scale);
    write_update(w, "Initializing Replicas", 1);
    ptx_engine_config cfg;
    ptx_engine_config_default(&cfg);
    cfg.chains = 2;
    cfg.sweeps = (int32_t)w.params.sweeps;
    cfg.flags = w.params.icm ? PTX_FLAG_ICM : 0u;
    cfg.nc = -1;
    const ptx_problem* probs[1] = {w.p.handle()};
    EngineHandle eng;
    check(ptx_engine_create(probs, 1, w.beta_values.data(), &nt, w.params.seed, &cfg, &eng.e));
    write_update(w, "Starting MCMC loop", 1);
    // progress in 2% chunks (solver.cc:108-126 prints a bar at verbosity >= 1)
    const uint64_t chunk = w.params.verbosity >= 1 ? std::max<uint64_t>(1, w.params.pts / 50) : w.params.pts;
    for (uint64_t done = 0; done < w.params.pts;) {
        const uint64_t n = std::min<uint64_t>(chunk, w.params.pts - done);
        check(ptx_engine_step(eng.e, (int64_t)n));
        done += n;
        if (w.params.verbosity >= 1) {
            check(ptx_engine_sync(eng.e));
            std::printf("\r%3llu%% Done", (unsigned long long)(done * 100 / std::max<uint64_t>(1, w.params.pts)));
            std::fflush(stdout);
        }
    }
    check(ptx_engine_sync(eng.e));
    if (w.params.verbosity >= 1) std::printf("\n");
    write_update(w, "Exiting MCMC loop", 1);
    // gather results (mpi_utils.cc:103-176 on one rank)
    ptx_instance_info ii;
    check(ptx_engine_instance_info(eng.e, 0, &ii));
    const size_t nrep = ii.nrep, n = ii.n;
    std::vector<int64_t> be(nrep), bs(nrep), E(nrep);
    std::vector<uint8_t> bst(nrep * n), st(nrep * n);
    std::vector<uint64_t> fl(nrep);
    std::vector<int32_t> s2r(nrep);
    check(ptx_engine_read_best(eng.e, 0, be.data(), bs.data(), bst.data()));
    check(ptx_engine_read_energies(eng.e, 0, E.data()));
    check(ptx_engine_read_spins(eng.e, 0, st.data()));
    check(ptx_engine_read_flips(eng.e, 0, fl.data()));
    check(ptx_engine_read_perm(eng.e, 0, s2r.data()));
    w.swap_attempts.assign(nrep, 0);
    w.swap_successes.assign(nrep, 0);
    uint64_t swaps = 0;
    check(ptx_engine_read_swap_stats(eng.e, 0, w.swap_attempts.data(), w.swap_successes.data(), &swaps));
    w.swaps = swaps;
    w.beta_indices.assign(s2r.begin(), s2r.end());
    w.replicas.assign(nrep, Replica());
    std::vector<int32_t> rep2slot(nrep);
    for (size_t s = 0; s < nrep; ++s) rep2slot[s2r[s]] = (int32_t)s;
    best_ene = 1.79769313486231570815e+308;
    size_t best_sweep = 0, best_r = 0;
    flips = 0;
    for (size_t r = 0; r < nrep; ++r) {
        Replica& R = w.replicas[r];
        R.energy = (double)E[r];
        R.beta = w.beta_values[rep2slot[r]];
        R.flips = fl[r];
        R.best_ene = be[r] == INT64_MAX ? 1.79769313486231570815e+308 : (double)be[r];
        R.best_sweep = (size_t)bs[r];
        R.state.assign(n, false);
        R.best_state.assign(n, false);
        for (size_t i = 0; i < n; ++i) {
            R.state[i] = st[r * n + i] != 0;
            R.best_state[i] = bst[r * n + i] != 0;
        }
        flips += fl[r];
        if (R.best_ene < best_ene) {
            best_ene = R.best_ene;
            best_sweep = R.best_sweep;
            best_r = r;
        }
    }
    best_state = w.replicas[best_r].best_state;
    w.elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    print_results(w, best_ene, best_sweep, flips, best_state);
}

void print_results(const PTWorker& w, double best_ene, size_t best_sweep, size_t flips, const StateVector& best_state) {
    const size_t nrep = w.swap_attempts.size();
    std::string line = "swap_rates: [";
    char buf[64];
    for (size_t i = 0; i < nrep; ++i) {  // utils.cc:4-10: 0 when never attempted
        const double r = w.swap_attempts[i] > 0 ? (double)w.swap_successes[i] / w.swap_attempts[i] : 0.0;
        std::snprintf(buf, sizeof buf, "%s%.6f", i ? ", " : "", r);
        line += buf;
    }
    std::printf("%s]\n", line.c_str());
    auto print_u32 = [](const char* key, const std::vector<uint32_t>& v) {
        std::printf("%s: [", key);
        for (size_t i = 0; i < v.size(); ++i) std::printf("%s%u", i ? ", " : "", v[i]);
        std::printf("]\n");
    };
    print_u32("swap_attempts", w.swap_attempts);
    print_u32("swap_successes", w.swap_successes);
    std::printf("best_state_energy: %s\n", fmt_double(w.p.energy(best_state)).c_str());
    std::printf("best_state_sweep: %zu\n", best_sweep);
    std::printf("best_energy: %s \n", fmt_double(best_ene).c_str());
    std::printf("best_maxcut: %s\n", fmt_double(w.p.maxcut(best_state)).c_str());
    std::printf("num_of_flips: %zu\n", flips);
    std::printf("num_of_swaps: %llu\n", (unsigned long long)w.swaps);
    std::printf("best_state: [");
    for (size_t i = 0; i < best_state.size(); ++i) std::printf("%s%d", i ? ", " : "", best_state[i] ? 1 : 0);
    std::printf("]\n");
    for (size_t i = 0; i < best_state.size(); ++i) std::putchar(best_state[i] ? '1' : '0');
    std::printf("\n[");
    const size_t nt = w.params.Nt;
    for (size_t i = 0; i < nt; ++i) std::printf("%s%.6f", i ? ", " : "", w.beta_values[i]);
    std::printf("]\n");
    std::printf("Total time: %.6f\n", w.elapsed);
    std::fflush(stdout);
}

}  // namespace pticm
