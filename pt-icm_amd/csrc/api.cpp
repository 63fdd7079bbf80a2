// api.cpp -- the extern "C" boundary declared in include/ptx.h.  Every entry
// point catches C++ exceptions, records a thread-local message and returns a
// status code; nothing here exits the process.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/ptx.h"
#include "engine.hpp"
#include "ladder.hpp"
#include "mlp.hpp"
#include "problem.hpp"

namespace ptx {
Status selftest_fp64(const double* x, int64_t n, double* eo, double* lo);
Status selftest_philox(const uint32_t* ctr, int64_t n, const uint32_t* key, uint32_t* out);
Status selftest_copy(int64_t nbytes, int32_t reps, double* ms_per_copy);
Status selftest_lds(int32_t blocks, int32_t iters, double* gbs);
}  // namespace ptx

struct ptx_problem {
    std::shared_ptr<const ptx::Problem> p;
};
struct ptx_engine {
    ptx::Engine e;
};
struct ptx_mlp {
    ptx::Mlp m;
};

namespace {

thread_local std::string g_last_error;

ptx_status fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
ptx_status done(const ptx::Status& s) {
    if (s) return PTX_OK;
    return fail(s.code, s.msg);
}

template <class F>
ptx_status guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return fail(PTX_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception& ex) {
        return fail(PTX_ERR_INVALID, ex.what());
    } catch (...) {
        return fail(PTX_ERR_INVALID, "unknown exception");
    }
}

}  // namespace

extern "C" {

ptx_status ptx_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return PTX_ERR_INVALID;
    std::strncpy(buf, g_last_error.c_str(), len - 1);
    buf[len - 1] = 0;
    return PTX_OK;
}

const char* ptx_version(void) { return "ptx 0.1 (gfx950)"; }

int32_t ptx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

ptx_status ptx_problem_read_gset(const char* path, ptx_problem** out) {
    return guarded([&]() -> ptx_status {
        if (!path || !out) return fail(PTX_ERR_INVALID, "null argument");
        auto P = std::make_shared<ptx::Problem>();
        ptx::Status s = ptx::read_gset(path, *P);
        if (!s) return done(s);
        *out = new ptx_problem{P};
        return PTX_OK;
    });
}

ptx_status ptx_problem_create_coo(int32_t n, int64_t m, const int32_t* src, const int32_t* dst, const double* w,
                                  const double* h, int32_t dense, ptx_problem** out) {
    return guarded([&]() -> ptx_status {
        if (!out) return fail(PTX_ERR_INVALID, "null argument");
        auto P = std::make_shared<ptx::Problem>();
        ptx::Status s = ptx::build_from_coo(n, m, src, dst, w, h, dense != 0, *P);
        if (!s) return done(s);
        *out = new ptx_problem{P};
        return PTX_OK;
    });
}

void ptx_problem_destroy(ptx_problem* p) { delete p; }

ptx_status ptx_problem_info_get(const ptx_problem* p, ptx_problem_info* out) {
    if (!p || !out) return fail(PTX_ERR_INVALID, "null argument");
    const ptx::Problem& P = *p->p;
    out->n = P.n;
    out->m = P.m_header;
    out->nnz = P.nnz();
    out->sparse = P.sparse ? 1 : 0;
    out->char_scale = P.char_scale;
    out->n_colours = P.ncls;
    out->fmax = P.fmax;
    out->e_min_bound = -P.abs_sum;
    out->e_max_bound = P.abs_sum;
    out->has_field = P.has_field ? 1 : 0;
    out->lattice_dim = P.lat_dim;
    out->lattice_L = P.lat_L;
    return PTX_OK;
}

ptx_status ptx_problem_energy(const ptx_problem* p, const uint8_t* state, int64_t* energy) {
    if (!p || !state || !energy) return fail(PTX_ERR_INVALID, "null argument");
    *energy = p->p->energy(state);
    return PTX_OK;
}

ptx_status ptx_problem_maxcut(const ptx_problem* p, const uint8_t* state, int64_t* cut) {
    if (!p || !state || !cut) return fail(PTX_ERR_INVALID, "null argument");
    *cut = p->p->maxcut(state);
    return PTX_OK;
}

ptx_status ptx_problem_delta_e(const ptx_problem* p, const uint8_t* state, int64_t* dE, int64_t* energy) {
    if (!p || !state || !dE) return fail(PTX_ERR_INVALID, "null argument");
    const int64_t e = p->p->delta_e(state, dE);
    if (energy) *energy = e;
    return PTX_OK;
}

ptx_status ptx_problem_colouring(const ptx_problem* p, int32_t* pos2orig, int32_t* class_start) {
    if (!p || !pos2orig || !class_start) return fail(PTX_ERR_INVALID, "null argument");
    const ptx::Problem& P = *p->p;
    std::memcpy(pos2orig, P.pos2orig.data(), sizeof(int32_t) * P.n);
    std::memcpy(class_start, P.class_start.data(), sizeof(int32_t) * (P.ncls + 1));
    return PTX_OK;
}

ptx_status ptx_problem_csr(const ptx_problem* p, int32_t* rowptr, int32_t* col, int32_t* w, int32_t* h) {
    if (!p) return fail(PTX_ERR_INVALID, "null argument");
    const ptx::Problem& P = *p->p;
    if (rowptr) std::memcpy(rowptr, P.rowptr.data(), sizeof(int32_t) * (P.n + 1));
    if (col) std::memcpy(col, P.col.data(), sizeof(int32_t) * P.nnz());
    if (w) std::memcpy(w, P.w.data(), sizeof(int32_t) * P.nnz());
    if (h) std::memcpy(h, P.h.data(), sizeof(int32_t) * P.n);
    return PTX_OK;
}

ptx_status ptx_beta_linspace(double* out, double T0, double T1, int64_t n, int32_t inverse) {
    if (!out || n < 1) return fail(PTX_ERR_INVALID, "bad linspace arguments");
    ptx::beta_linspace(out, T0, T1, n, inverse != 0);
    return PTX_OK;
}

int32_t ptx_icm_count(const double* beta, int32_t nt, double char_scale) {
    if (!beta || nt < 0) return 0;
    return ptx::icm_count(beta, nt, char_scale);
}

ptx_status ptx_pilot_betas(double T0, double T1, int32_t nt, int32_t inverse, double* out, double* bs, double* be) {
    if (!out || !bs || !be || nt < 2) return fail(PTX_ERR_INVALID, "the pilot ladder needs at least 2 temperatures");
    ptx::pilot_betas(T0, T1, nt, inverse != 0, out, bs, be);
    return PTX_OK;
}

ptx_status ptx_ladder_from_histograms(const uint32_t* counts, int32_t nt, int64_t nbins, int64_t e_base,
                                      const double* hist_betas, double estep, double step_size, double r,
                                      double beta_start, double beta_end, int32_t mode, double* betas_out,
                                      int32_t* nt_out) {
    return guarded([&]() -> ptx_status {
        if (!counts || !hist_betas || !betas_out || !nt_out || nt < 2 || nbins < 1 || !(estep > 0))
            return fail(PTX_ERR_INVALID, "bad ladder arguments");
        auto H = ptx::rebin(counts, nt, nbins, e_base, estep);
        for (const auto& s : H)
            if (s.h.empty()) return fail(PTX_ERR_INVALID, "a pilot histogram is empty");
        ptx::Dos D = mode == PTX_LADDER_NEWTON ? ptx::solve_dos_newton(H, hist_betas, estep)
                                               : ptx::solve_dos(H, hist_betas, estep);
        *nt_out = ptx::ladder_search(D, nt, beta_start, beta_end, step_size, r, mode != PTX_LADDER_EXACT, betas_out);
        return PTX_OK;
    });
}

ptx_status ptx_ladder_predict_rates(const uint32_t* counts, int32_t nt, int64_t nbins, int64_t e_base,
                                    const double* hist_betas, double estep, const double* ladder, int32_t n_ladder,
                                    int32_t mode, double* rates_out) {
    return guarded([&]() -> ptx_status {
        if (!counts || !hist_betas || !ladder || !rates_out || nt < 2 || n_ladder < 2)
            return fail(PTX_ERR_INVALID, "bad arguments");
        auto H = ptx::rebin(counts, nt, nbins, e_base, estep);
        for (const auto& s : H)
            if (s.h.empty()) return fail(PTX_ERR_INVALID, "a pilot histogram is empty");
        ptx::Dos D = mode == PTX_LADDER_NEWTON ? ptx::solve_dos_newton(H, hist_betas, estep)
                                               : ptx::solve_dos(H, hist_betas, estep);
        std::vector<double> p1, p2;
        for (int32_t k = 0; k + 1 < n_ladder; ++k) {
            double b1 = ladder[k], b2 = ladder[k + 1];
            if (b1 < b2) std::swap(b1, b2);
            const double z1 = ptx::partition_func(D.logdos, b1, D.min_emin, estep);
            const double z2 = ptx::partition_func(D.logdos, b2, D.min_emin, estep);
            p1.resize(D.logdos.size());
            p2.resize(D.logdos.size());
            double e = D.min_emin;
            for (size_t i = 0; i < D.logdos.size(); ++i, e += estep) {
                p1[i] = D.logdos[i] - e * b1 - z1;
                p2[i] = D.logdos[i] - e * b2 - z2;
            }
            rates_out[k] = ptx::acceptance_rate(p1, p2, b1, b2, estep, D.min_emin, mode != PTX_LADDER_EXACT);
        }
        return PTX_OK;
    });
}

void ptx_engine_config_default(ptx_engine_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->chains = 2;
    c->sweeps = 1;
    c->flags = 0;
    c->nc = -1;
    c->hist_start = 0;
    c->mode = PTX_MODE_AUTO;
    c->waves_per_group = 0;
    c->steps_per_launch = 0;
    c->use_graph = 0;
    c->spins_per_lane = 0;
    c->csr_placement = PTX_CSR_AUTO;
    c->icm_waves = 0;
    c->small_batch_coop = 0;
    c->coop_wait_us = 0;
    c->sliced_sweep = 0;
}

ptx_status ptx_engine_create(const ptx_problem* const* problems, int32_t n_instances, const double* betas,
                             const int32_t* n_temps, uint64_t seed, const ptx_engine_config* cfg, ptx_engine** out) {
    return guarded([&]() -> ptx_status {
        if (!problems || !betas || !n_temps || !out || n_instances <= 0) return fail(PTX_ERR_INVALID, "null argument");
        ptx_engine_config c;
        if (cfg) c = *cfg;
        else ptx_engine_config_default(&c);
        std::vector<std::shared_ptr<const ptx::Problem>> ps(n_instances);
        for (int32_t i = 0; i < n_instances; ++i) {
            if (!problems[i]) return fail(PTX_ERR_INVALID, "null problem");
            ps[i] = problems[i]->p;
        }
        std::unique_ptr<ptx_engine> e(new ptx_engine());
        ptx::Status s = e->e.create(ps, betas, n_temps, seed, c);
        if (!s) return done(s);
        *out = e.release();
        return PTX_OK;
    });
}

void ptx_engine_destroy(ptx_engine* e) { delete e; }

ptx_status ptx_engine_autotune(const ptx_problem* const* problems, int32_t n_instances, const double* betas,
                               const int32_t* n_temps, uint64_t seed, const ptx_engine_config* base, int32_t steps,
                               ptx_engine_config* best, ptx_tune_trial* trials, int32_t max_trials,
                               int32_t* n_trials) {
    return guarded([&]() -> ptx_status {
        if (!problems || !betas || !n_temps || !best || n_instances <= 0 || steps <= 0)
            return fail(PTX_ERR_INVALID, "bad autotune arguments");
        ptx_engine_config b0;
        if (base) b0 = *base;
        else ptx_engine_config_default(&b0);
        std::vector<std::shared_ptr<const ptx::Problem>> ps(n_instances);
        for (int32_t i = 0; i < n_instances; ++i) {
            if (!problems[i]) return fail(PTX_ERR_INVALID, "null problem");
            ps[i] = problems[i]->p;
        }
        struct Shape {
            int32_t mode, w, spl, csr, wi;
        };
        std::vector<Shape> shapes;
        if (n_instances == 1) {
            for (int w : {4, 8})
                for (int spl : {1, 4}) shapes.push_back({PTX_MODE_COOP, w, spl, PTX_CSR_AUTO, 0});
            shapes.push_back({PTX_MODE_COOP, 16, 1, PTX_CSR_AUTO, 0});  // taken as 8 with cluster moves
        }
        // fused: 4 or 8 waves (the 16-wave instantiation spills registers to
        // scratch and writes many times its state per launch: not offered)
        for (int w : {4, 8})
            for (int csr : {PTX_CSR_LDS, PTX_CSR_GLOBAL}) shapes.push_back({PTX_MODE_FUSED, w, 0, csr, 0});
        // split: one wave per cluster (the default, 0) or four
        const bool icm = (b0.flags & PTX_FLAG_ICM) != 0;
        for (int w : {1, 4})
            for (int wi : {0, 4}) {
                if (!icm && wi != 0) continue;  // no cluster moves: one split shape per sweep width
                shapes.push_back({PTX_MODE_SPLIT, w, 0, PTX_CSR_AUTO, wi});
            }
        double best_us = 0;
        int32_t cnt = 0;
        bool found = false;
        for (const Shape& sh : shapes) {
            ptx_engine_config c = b0;
            c.mode = sh.mode;
            c.waves_per_group = sh.w;
            c.spins_per_lane = sh.spl;
            c.csr_placement = sh.csr;
            c.icm_waves = sh.wi;
            ptx::Engine e;
            if (!e.create(ps, betas, n_temps, seed, c)) continue;  // shape does not fit this batch
            if (e.mode != sh.mode) continue;
            const int64_t warm = e.mode == PTX_MODE_SPLIT ? 2 : e.K;
            if (!e.step(warm) || !e.sync()) continue;
            const auto t0 = std::chrono::steady_clock::now();
            if (!e.step(steps) || !e.sync()) continue;
            const double us =
                std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count() / steps;
            if (trials && cnt < max_trials)
                trials[cnt] = ptx_tune_trial{sh.mode, e.mode == PTX_MODE_SPLIT ? e.W : (e.mode == PTX_MODE_COOP ? e.coop_W : e.Wf), sh.spl, sh.csr,
                                             e.mode == PTX_MODE_SPLIT ? e.icm_waves_now() : 0, us};
            ++cnt;
            if (!found || us < best_us) {
                found = true;
                best_us = us;
                *best = c;
            }
        }
        if (n_trials) *n_trials = cnt;
        if (!found) return fail(PTX_ERR_UNSUPPORTED, "no launch shape fits this batch");
        return PTX_OK;
    });
}

ptx_status ptx_engine_step(ptx_engine* e, int64_t n_steps) {
    if (!e) return fail(PTX_ERR_INVALID, "null engine");
    return guarded([&]() { return done(e->e.step(n_steps)); });
}

ptx_status ptx_engine_set_active(ptx_engine* e, const uint8_t* active) {
    if (!e || !active) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() { return done(e->e.set_active(active)); });
}

ptx_status ptx_engine_sync(ptx_engine* e) {
    if (!e) return fail(PTX_ERR_INVALID, "null engine");
    return done(e->e.sync());
}

ptx_status ptx_engine_info_get(const ptx_engine* e, ptx_engine_info* out) {
    if (!e || !out) return fail(PTX_ERR_INVALID, "null argument");
    const ptx::Engine& E = e->e;
    out->n_instances = E.ninst;
    out->steps_done = E.steps_done;
    out->mode = E.last_mode;
    out->waves_per_group = E.last_mode == PTX_MODE_SPLIT ? E.W : (E.last_mode == PTX_MODE_COOP ? E.coop_W : E.Wf);
    out->steps_per_launch = E.last_mode == PTX_MODE_SPLIT ? 1 : (E.last_mode == PTX_MODE_COOP ? E.coop_K : E.K);
    out->total_replicas = E.total_rep;
    out->spin_updates_per_step = E.updates_per_step;
    out->lds_bytes_per_step = E.lds_bytes_per_step;
    out->csr_in_lds = E.last_mode == PTX_MODE_COOP ? (E.coop_last_form == 0 ? 1 : 0)
                      : (E.last_mode == PTX_MODE_SPLIT ? 0 : (E.fused_global ? 0 : 1));
    out->icm_waves = E.last_mode == PTX_MODE_SPLIT ? E.icm_waves_now() : 1;
    out->coop_capacity = E.coop_capacity;
    return PTX_OK;
}

ptx_status ptx_engine_instance_info(const ptx_engine* e, int32_t inst, ptx_instance_info* out) {
    if (!e || !out || inst < 0 || inst >= e->e.ninst) return fail(PTX_ERR_INVALID, "bad instance");
    const ptx::InstDesc& d = e->e.desc[inst];
    out->n = d.n;
    out->nt = d.nt;
    out->nrep = d.nrep;
    out->nc = d.nc;
    out->n_colours = d.ncls;
    out->hist_bins = d.hist_bins;
    out->e_base = d.e_base;
    out->lattice_sweep = d.lsw_dim;
    out->sliced_sweep = d.rs;
    return PTX_OK;
}

#define PTX_ENGINE_CALL(expr)                                   \
    do {                                                        \
        if (!e) return fail(PTX_ERR_INVALID, "null engine");    \
        return guarded([&]() { return done(expr); });           \
    } while (0)

ptx_status ptx_engine_read_spins(ptx_engine* e, int32_t inst, uint8_t* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_spins(inst, out));
}
ptx_status ptx_engine_write_spins(ptx_engine* e, int32_t inst, const uint8_t* in) {
    if (!in) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.write_spins(inst, in));
}
ptx_status ptx_engine_read_energies(ptx_engine* e, int32_t inst, int64_t* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_energies(inst, out));
}
ptx_status ptx_engine_read_perm(ptx_engine* e, int32_t inst, int32_t* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_perm(inst, out));
}
ptx_status ptx_engine_read_swap_stats(ptx_engine* e, int32_t inst, uint32_t* att, uint32_t* succ, uint64_t* swaps) {
    if (!att || !succ || !swaps) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_swap_stats(inst, att, succ, swaps));
}
ptx_status ptx_engine_read_flips(ptx_engine* e, int32_t inst, uint64_t* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_flips(inst, out));
}
ptx_status ptx_engine_read_best(ptx_engine* e, int32_t inst, int64_t* be, int64_t* bs, uint8_t* st) {
    if (!be || !bs) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_best(inst, be, bs, st));
}
ptx_status ptx_engine_read_histograms(ptx_engine* e, int32_t inst, uint32_t* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_histograms(inst, out));
}
ptx_status ptx_engine_read_observables(ptx_engine* e, int32_t inst, ptx_slot_observables* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_observables(inst, out));
}
ptx_status ptx_engine_state_size(const ptx_engine* e, int64_t* bytes) {
    if (!e || !bytes) return fail(PTX_ERR_INVALID, "null argument");
    *bytes = e->e.state_bytes();
    return PTX_OK;
}
ptx_status ptx_engine_save_state(ptx_engine* e, void* buf, int64_t bytes) {
    if (!buf) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.save_state(buf, bytes));
}
ptx_status ptx_engine_load_state(ptx_engine* e, const void* buf, int64_t bytes) {
    if (!buf) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.load_state(buf, bytes));
}
ptx_status ptx_engine_read_min_energies(ptx_engine* e, int64_t* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.read_min_energies(out));
}
ptx_status ptx_engine_timing_enable(ptx_engine* e, int32_t on) {
    if (!e) return fail(PTX_ERR_INVALID, "null engine");
    e->e.timing = on != 0;
    return PTX_OK;
}
ptx_status ptx_engine_timing_get(ptx_engine* e, ptx_engine_timing* out) {
    if (!out) return fail(PTX_ERR_INVALID, "null buffer");
    PTX_ENGINE_CALL(e->e.timing_get(out));
}

ptx_status ptx_engine_phase_profile(ptx_engine* e, int32_t on, double* us) {
    PTX_ENGINE_CALL(e->e.phase_profile(on, us));
}

ptx_status ptx_selftest_fp64(const double* x, int64_t n, double* exp_out, double* log_out) {
    if (!x || !exp_out || !log_out || n < 0) return fail(PTX_ERR_INVALID, "null buffer");
    return guarded([&]() { return done(ptx::selftest_fp64(x, n, exp_out, log_out)); });
}

ptx_status ptx_host_fp64(const double* x, int64_t n, double* exp_out, double* log_out) {
    if (!x || !exp_out || !log_out || n < 0) return fail(PTX_ERR_INVALID, "null buffer");
    for (int64_t i = 0; i < n; ++i) {
        exp_out[i] = ptx::pexp(x[i]);
        log_out[i] = ptx::plog(x[i]);
    }
    return PTX_OK;
}

ptx_status ptx_selftest_copy(int64_t nbytes, int32_t reps, double* ms_per_copy) {
    if (!ms_per_copy) return fail(PTX_ERR_INVALID, "null buffer");
    return guarded([&]() { return done(ptx::selftest_copy(nbytes, reps, ms_per_copy)); });
}

ptx_status ptx_selftest_lds_bandwidth(int32_t blocks, int32_t iters, double* gbs) {
    if (!gbs) return fail(PTX_ERR_INVALID, "null buffer");
    return guarded([&]() { return done(ptx::selftest_lds(blocks, iters, gbs)); });
}

ptx_status ptx_selftest_philox(const uint32_t* ctr, int64_t n, const uint32_t* key, uint32_t* out) {
    if (!ctr || !key || !out || n < 0) return fail(PTX_ERR_INVALID, "null buffer");
    return guarded([&]() { return done(ptx::selftest_philox(ctr, n, key, out)); });
}

ptx_status ptx_mlp_create(const int32_t* dims, int32_t n_dims, int32_t max_batch, uint64_t seed, ptx_mlp** out) {
    if (!dims || !out) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() -> ptx_status {
        std::unique_ptr<ptx_mlp> h(new ptx_mlp);
        ptx::Status s = h->m.create(dims, n_dims, max_batch, seed);
        if (!s) return done(s);
        *out = h.release();
        return PTX_OK;
    });
}

void ptx_mlp_destroy(ptx_mlp* m) { delete m; }

ptx_status ptx_mlp_param_count(const ptx_mlp* m, int64_t* count) {
    if (!m || !count) return fail(PTX_ERR_INVALID, "null argument");
    *count = m->m.nparams;
    return PTX_OK;
}

ptx_status ptx_mlp_set_params(ptx_mlp* m, const float* params, int64_t count) {
    if (!m || !params) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() { return done(m->m.set_params(params, count)); });
}

ptx_status ptx_mlp_get_params(ptx_mlp* m, float* params, int64_t count) {
    if (!m || !params) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() { return done(m->m.get_params(params, count)); });
}

ptx_status ptx_mlp_get_grads(ptx_mlp* m, float* grads, int64_t count) {
    if (!m || !grads) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() { return done(m->m.get_grads(grads, count)); });
}

ptx_status ptx_mlp_forward(ptx_mlp* m, const float* x, int32_t batch, float* y) {
    if (!m || !x || !y) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() { return done(m->m.forward(x, batch, y)); });
}

ptx_status ptx_mlp_forward_noalloc(const int32_t* dims, int32_t n_dims, const float* params, int64_t param_count,
                                   const float* x, int32_t batch, float* y) {
    if (!dims || !params || !x || !y) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() { return done(ptx::mlp_forward_noalloc(dims, n_dims, params, param_count, x, batch, y)); });
}

ptx_status ptx_mlp_loss_grad(ptx_mlp* m, const float* x, const float* target, const uint8_t* mode, int32_t batch,
                             double* loss) {
    if (!m || !x || !target || !mode) return fail(PTX_ERR_INVALID, "null argument");
    return guarded([&]() { return done(m->m.loss_grad(x, target, mode, batch, loss)); });
}

ptx_status ptx_mlp_adam(ptx_mlp* m, float lr, float beta1, float beta2, float eps) {
    if (!m) return fail(PTX_ERR_INVALID, "null argument");
    if (!(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return fail(PTX_ERR_INVALID, "Adam hyper-parameters out of range");
    return guarded([&]() { return done(m->m.adam(lr, beta1, beta2, eps)); });
}

ptx_status ptx_mlp_train(ptx_mlp* m, const float* x, const float* target, const uint8_t* mode, int64_t n_samples,
                         int32_t batch, int32_t epochs, float lr, uint64_t seed, double* epoch_loss) {
    if (!m || !x || !target || !mode) return fail(PTX_ERR_INVALID, "null argument");
    if (!(lr > 0.0f)) return fail(PTX_ERR_INVALID, "learning rate must be positive");
    return guarded([&]() { return done(m->m.train(x, target, mode, n_samples, batch, epochs, lr, seed, epoch_loss)); });
}

ptx_status ptx_selftest_gemm_bf16(int32_t M, int32_t N, int32_t K, const uint16_t* A, const uint16_t* B, float* C) {
    if (!A || !B || !C) return fail(PTX_ERR_INVALID, "null buffer");
    return guarded([&]() { return done(ptx::selftest_gemm_bf16(M, N, K, A, B, C)); });
}

}  // extern "C"
