// pticm_facade.cpp -- see pticm.hh.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>

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

std::vector<double> read_ladder(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw Error(PTX_ERR_IO, "cannot read ladder file " + path);
    std::vector<double> b;
    double x;
    while (f >> x) b.push_back(x);
    if (b.empty()) throw Error(PTX_ERR_INVALID, "ladder file " + path + " holds no inverse temperature");
    return b;
}

void write_ladder(const std::string& path, const double* b, size_t n) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw Error(PTX_ERR_IO, "cannot write " + path);
    for (size_t i = 0; i < n; ++i) std::fprintf(f, "%.17g\n", b[i]);
    std::fclose(f);
}

std::vector<char> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error(PTX_ERR_IO, "cannot read " + path);
    return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

void write_file(const std::string& path, const void* data, size_t n) {
    std::ofstream f(path, std::ios::binary);
    if (!f || !f.write(static_cast<const char*>(data), (std::streamsize)n)) throw Error(PTX_ERR_IO, "cannot write " + path);
}

// The legacy observable writer of src/main.cc:493-528: per temperature
// T = 1/beta, C = beta^2 var(E) / n, X = beta n var(m), u = <E>/n and m =
// <up spins>/n, from the engine's per-slot sums pooled over both chains.
void write_observables(const std::string& path, ptx_engine* e, const std::vector<double>& beta, int32_t nt, int64_t n) {
    ptx_instance_info ii;
    check(ptx_engine_instance_info(e, 0, &ii));
    std::vector<ptx_slot_observables> o((size_t)ii.nrep);
    check(ptx_engine_read_observables(e, 0, o.data()));
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) throw Error(PTX_ERR_IO, "cannot write " + path);
    std::fprintf(f, "T,C,X,u,m\n");
    for (int32_t j = 0; j < nt; ++j) {
        double k = 0, se = 0, se2 = 0, su = 0, su2 = 0;
        for (int32_t c = 0; c < ii.nrep / nt; ++c) {
            const ptx_slot_observables& x = o[(size_t)c * nt + j];
            k += (double)x.samples;
            se += (double)x.sum_e;
            se2 += x.sum_e2;
            su += (double)x.sum_up;
            su2 += x.sum_up2;
        }
        const double b = beta[j], e1 = se / k, e2 = se2 / k, m1 = su / k / n, m2 = su2 / k / ((double)n * n);
        std::fprintf(f, "%f,%f,%f,%f,%f\n", 1.0 / b, b * b / n * (e2 - e1 * e1), b * n * (m2 - m1 * m1), e1 / n, m1);
    }
    std::fclose(f);
}

// std::vector printed as "[a, b, c]" with the stream's default formatting (src/utils.hh:12-23)
template <class T>
std::string list_str(const std::vector<T>& v) {
    std::ostringstream o;
    o << "[";
    for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
    o << "]";
    return o.str();
}

std::atomic<uint64_t> g_replica_serial{0};

}  // namespace

Problem::Problem(const std::string& path, uint64_t seed, uint64_t burn_in) : seed_(seed), burn_in_(burn_in) {
    ptx_problem* p = nullptr;
    check(ptx_problem_read_gset(path.c_str(), &p));
    h_.reset(p, ptx_problem_destroy);
    init_info();
}

Problem::Problem(double* formulation, double* linear, uint64_t seed, uint64_t burn_in, uint64_t N)
    : seed_(seed), burn_in_(burn_in) {
    std::vector<int32_t> src, dst;
    std::vector<double> w;
    for (uint64_t i = 0; i < N; ++i)
        for (uint64_t j = i + 1; j < N; ++j)
            if (formulation[i * N + j] != 0.0) {
                src.push_back((int32_t)i);
                dst.push_back((int32_t)j);
                w.push_back(formulation[i * N + j]);
            }
    ptx_problem* p = nullptr;
    check(ptx_problem_create_coo((int32_t)N, (int64_t)w.size(), src.data(), dst.data(), w.data(), linear, 1, &p));
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

void Problem::init_info() {
    check(ptx_problem_info_get(h_.get(), &info_));
    rng_.seed(seed_);
    auto c = std::make_shared<Csr>();
    c->rowptr.resize((size_t)info_.n + 1);
    c->col.resize((size_t)std::max<int64_t>(info_.nnz, 1));
    c->w.resize(c->col.size());
    c->h.resize((size_t)info_.n);
    check(ptx_problem_csr(h_.get(), c->rowptr.data(), c->col.data(), c->w.data(), c->h.data()));
    csr_ = c;
}

double Problem::energy(bool* s) const {
    std::vector<uint8_t> b(info_.n);
    for (int32_t i = 0; i < info_.n; ++i) b[i] = s[i] ? 1 : 0;
    int64_t e = 0;
    check(ptx_problem_energy(h_.get(), b.data(), &e));
    return (double)e;
}

void Problem::incremental_dE(std::vector<double>& dE, const StateVector& s, uint32_t index) const {
    dE.at(index) *= -1;
    const double si = s[index] ? 1.0 : -1.0;
    for (int32_t k = csr_->rowptr[index]; k < csr_->rowptr[index + 1]; ++k) {
        const int32_t j = csr_->col[k];
        dE.at(j) += 4.0 * si * (s[j] ? 1.0 : -1.0) * csr_->w[k];
    }
}

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

// ---- Replica (src/replica.cc:5-57) ----
struct Replica::Device {
    ptx_engine* e = nullptr;
    double beta = 0;
    size_t base_sweep = 0;   // the replica's sweep counter when the engine was made
    uint64_t flips_seen = 0; // the engine's flip count already added
    ~Device() { ptx_engine_destroy(e); }
};

Replica::Replica(Problem& p, double b, const StateVector& init_state, const std::vector<double>& init_dE, double e,
                 bool log_ene, size_t burn_in)
    : energy(e), beta(b), state(init_state), best_state(init_state), dE(init_dE), log_energy_(log_ene),
      burn_in_(burn_in) {
    (void)p;
    serial_ = g_replica_serial.fetch_add(1);
}

Replica& Replica::operator=(const Replica& o) {
    if (this == &o) return *this;
    energy = o.energy;
    beta = o.beta;
    best_ene = o.best_ene;
    flips = o.flips;
    best_sweep = o.best_sweep;
    state = o.state;
    best_state = o.best_state;
    dE = o.dE;
    log_energy_ = o.log_energy_;
    burn_in_ = o.burn_in_;
    sweep_counter_ = o.sweep_counter_;
    ene_log = o.ene_log;
    serial_ = g_replica_serial.fetch_add(1);  // a copy continues on streams of its own
    dev_.reset();
    dirty_ = true;
    return *this;
}

void Replica::sweep(Problem& p) {
    const int32_t n = (int32_t)p.vcount();
    if (!dev_ || dev_->beta != beta || dirty_) {
        auto d = std::make_shared<Device>();
        ptx_engine_config cfg;
        ptx_engine_config_default(&cfg);
        cfg.chains = 1;
        cfg.flags = PTX_FLAG_NO_EXCHANGE;
        cfg.nc = 0;
        const ptx_problem* probs[1] = {p.handle()};
        const int32_t nt = 1;
        const uint64_t seed = p.get_seed() ^ (0x9E3779B97F4A7C15ull * (serial_ + 1));
        check(ptx_engine_create(probs, 1, &beta, &nt, seed, &cfg, &d->e));
        std::vector<uint8_t> b(n);
        for (int32_t i = 0; i < n; ++i) b[i] = state[i] ? 1 : 0;
        check(ptx_engine_write_spins(d->e, 0, b.data()));
        d->beta = beta;
        d->base_sweep = sweep_counter_;
        dev_ = d;
        dirty_ = false;
    }
    check(ptx_engine_step(dev_->e, 1));
    std::vector<uint8_t> b(n), bst(n);
    int64_t e = 0, be = 0, bs = 0;
    uint64_t f = 0;
    check(ptx_engine_read_spins(dev_->e, 0, b.data()));
    check(ptx_engine_read_energies(dev_->e, 0, &e));
    check(ptx_engine_read_flips(dev_->e, 0, &f));
    check(ptx_engine_read_best(dev_->e, 0, &be, &bs, bst.data()));
    for (int32_t i = 0; i < n; ++i) state[i] = b[i] != 0;
    energy = (double)e;
    flips += f - dev_->flips_seen;
    dev_->flips_seen = f;
    if (be != INT64_MAX && (double)be < best_ene) {  // replica.cc:23-27 (checked per colour class)
        best_ene = (double)be;
        best_sweep = dev_->base_sweep + (size_t)bs;
        for (int32_t i = 0; i < n; ++i) best_state[i] = bst[i] != 0;
    }
    p.calculate_dE(dE, state);
    if (log_energy_ && sweep_counter_ > burn_in_) ene_log.push_back(energy);
    ++sweep_counter_;
}

// ---- ladders ----
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
    if (!w.params.histograms_out.empty()) {  // slot, its beta, energy, count (nonzero counts)
        FILE* f = std::fopen(w.params.histograms_out.c_str(), "w");
        if (!f) throw Error(PTX_ERR_IO, "cannot write " + w.params.histograms_out);
        std::fprintf(f, "slot,beta,energy,count\n");
        for (int32_t j = 0; j < nt; ++j)
            for (int32_t k = 0; k < ii.hist_bins; ++k)
                if (counts[(size_t)j * ii.hist_bins + k])
                    std::fprintf(f, "%d,%.17g,%lld,%u\n", j, hist_betas[j], (long long)(ii.e_base + 2 * k),
                                 counts[(size_t)j * ii.hist_bins + k]);
        std::fclose(f);
    }
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

// The temperature ladder of one solve: from a file, the preprocess, or linear in beta.
static void make_ladder(PTWorker& w) {
    if (!w.params.ladder_in.empty()) {
        const std::vector<double> b = read_ladder(w.params.ladder_in);
        w.params.Nt = b.size();
        w.beta_values.assign(2 * b.size(), 0.0);
        std::copy(b.begin(), b.end(), w.beta_values.begin());
    } else if (w.params.preprocess) {
        get_beta(w, 2, w.params.pre_step, 0.4, w.params.pre_samples);
    } else {
        w.beta_values.assign(2 * w.params.Nt, 0.0);
        beta_linspace(w.beta_values.data(), w.params.T0, w.params.T1, w.params.Nt, w.params.inverse);
    }
    const size_t nt = w.params.Nt;
    for (size_t i = 0; i < nt; ++i) w.beta_values[nt + i] = w.beta_values[i];
    if (!w.params.ladder_out.empty()) write_ladder(w.params.ladder_out, w.beta_values.data(), nt);
}

static ptx_engine_config main_config(const Parameters& par) {
    ptx_engine_config cfg;
    ptx_engine_config_default(&cfg);
    cfg.chains = 2;
    cfg.sweeps = (int32_t)par.sweeps;
    cfg.flags = (par.icm ? PTX_FLAG_ICM : 0u) | (par.log_data ? PTX_FLAG_OBSERVE : 0u);
    cfg.hist_start = par.log_data ? (int64_t)par.burn_in : 0;  // observables after the burn-in
    cfg.nc = -1;
    return cfg;
}

void solve(PTWorker& w, std::string& problem_path, double& best_ene, size_t& flips, StateVector& best_state) {
    const auto t0 = std::chrono::steady_clock::now();
    w.p = Problem(problem_path, w.params.seed, w.params.burn_in);
    w.characteristic_scale = w.p.get_char_scale();
    write_update(w, "Initializing Temperatures", 1);
    make_ladder(w);
    const int32_t nt = (int32_t)w.params.Nt;
    w.Nc = (uint32_t)ptx_icm_count(w.beta_values.data(), nt, w.characteristic_scale);
    write_update(w, "Initializing Replicas", 1);
    const ptx_engine_config cfg = main_config(w.params);
    const ptx_problem* probs[1] = {w.p.handle()};
    EngineHandle eng;
    check(ptx_engine_create(probs, 1, w.beta_values.data(), &nt, w.params.seed, &cfg, &eng.e));
    uint64_t done = 0;
    if (!w.params.resume.empty()) {  // continue a run stopped with --checkpoint
        const std::vector<char> blob = read_file(w.params.resume);
        check(ptx_engine_load_state(eng.e, blob.data(), (int64_t)blob.size()));
        ptx_engine_info info;
        check(ptx_engine_info_get(eng.e, &info));
        done = (uint64_t)info.steps_done;
    }
    write_update(w, "Starting MCMC loop", 1);
    // progress in 2% chunks (solver.cc:108-126 prints a bar at verbosity >= 1)
    const uint64_t chunk = w.params.verbosity >= 1 ? std::max<uint64_t>(1, w.params.pts / 50) : w.params.pts;
    while (done < w.params.pts) {
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
    if (!w.params.checkpoint.empty()) {
        int64_t bytes = 0;
        check(ptx_engine_state_size(eng.e, &bytes));
        std::vector<char> blob((size_t)bytes);
        check(ptx_engine_save_state(eng.e, blob.data(), bytes));
        write_file(w.params.checkpoint, blob.data(), blob.size());
    }
    if (w.params.log_data) write_observables(w.params.log_path, eng.e, w.beta_values, nt, (int64_t)w.p.vcount());
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

void solve_folder(PTWorker& w, const std::string& folder, const std::string& csv, size_t restarts) {
    std::vector<std::string> paths;
    for (const auto& entry : std::filesystem::directory_iterator(folder))
        if (entry.is_regular_file()) paths.push_back(entry.path().string());
    std::sort(paths.begin(), paths.end());
    if (paths.empty()) throw Error(PTX_ERR_INVALID, "no problem files in " + folder);
    std::ofstream out(csv);
    if (!out) throw Error(PTX_ERR_IO, "cannot write " + csv);
    out << "problem,iter,pts,replicas,tlist,pswap,icm,preprocess,burn_in,inverse,sweeps,ene,state,cut,flips\n";
    const Parameters base = w.params;
    const size_t np = paths.size();
    for (size_t it = 0; it < restarts; ++it) {
        // every problem of the folder, with its own ladder, in one batched engine
        std::vector<PTWorker> ws(np);
        std::vector<const ptx_problem*> probs(np);
        std::vector<double> betas;
        std::vector<int32_t> nts(np);
        for (size_t k = 0; k < np; ++k) {
            ws[k].params = base;
            ws[k].params.seed = base.seed + it;  // main.cc:203 (seed++ per restart)
            ws[k].params.ladder_out.clear();
            ws[k].params.histograms_out.clear();
            ws[k].p = Problem(paths[k], ws[k].params.seed, base.burn_in);
            make_ladder(ws[k]);
            probs[k] = ws[k].p.handle();
            nts[k] = (int32_t)ws[k].params.Nt;
            betas.insert(betas.end(), ws[k].beta_values.begin(), ws[k].beta_values.begin() + nts[k]);
        }
        ptx_engine_config cfg = main_config(base);
        cfg.flags &= ~PTX_FLAG_OBSERVE;
        EngineHandle eng;
        check(ptx_engine_create(probs.data(), (int32_t)np, betas.data(), nts.data(), base.seed + it, &cfg, &eng.e));
        check(ptx_engine_step(eng.e, (int64_t)base.pts));
        check(ptx_engine_sync(eng.e));
        for (size_t k = 0; k < np; ++k) {
            const size_t nrep = 2 * (size_t)nts[k], n = ws[k].p.vcount();
            std::vector<int64_t> be(nrep), bs(nrep);
            std::vector<uint8_t> bst(nrep * n);
            std::vector<uint64_t> fl(nrep);
            std::vector<uint32_t> att(nrep), suc(nrep);
            uint64_t swaps = 0;
            check(ptx_engine_read_best(eng.e, (int32_t)k, be.data(), bs.data(), bst.data()));
            check(ptx_engine_read_flips(eng.e, (int32_t)k, fl.data()));
            check(ptx_engine_read_swap_stats(eng.e, (int32_t)k, att.data(), suc.data(), &swaps));
            const size_t r = (size_t)(std::min_element(be.begin(), be.end()) - be.begin());
            StateVector best(n);
            for (size_t i = 0; i < n; ++i) best[i] = bst[r * n + i] != 0;
            std::vector<double> pswap(nrep);
            for (size_t s = 0; s < nrep; ++s) pswap[s] = (double)suc[s] / (double)att[s];  // main.cc:181
            uint64_t nfl = 0;
            for (uint64_t x : fl) nfl += x;
            std::vector<int> bits(best.begin(), best.end());
            // list columns are quoted so the row parses as CSV (the reference writes them bare)
            out << paths[k] << "," << it << "," << base.pts << "," << nts[k] << ",\""
                << list_str(std::vector<double>(ws[k].beta_values.begin(), ws[k].beta_values.begin() + nrep))
                << "\",\"" << list_str(pswap) << "\"," << base.icm << "," << base.preprocess << "," << base.burn_in
                << "," << base.inverse << "," << base.sweeps << "," << (double)be[r] << ",\"" << list_str(bits)
                << "\"," << ws[k].p.maxcut(best) << "," << nfl << "\n";
        }
    }
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
