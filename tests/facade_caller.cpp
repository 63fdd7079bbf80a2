// A caller written against the reference's C++ interface (namespace pticm:
// Problem, Replica, beta_linspace), compiled against the facade
// pt-icm_amd/csrc/cli/pticm.hh instead.  It follows the reference's own
// initialisation (src/initialization.cc:63-76: Bernoulli states from the
// problem's random stream, calculate_dE, Replica(p, beta, state, dE, E)) and
// drives the replicas' sweeps itself.
//
//   facade_caller host GRAPH   checks that need no GPU: the incremental dE
//                              update (problem.cc:164-181) against a full
//                              recompute after every single flip, energy of
//                              bool arrays, the host random stream
//   facade_caller gpu GRAPH    sweeps of several replicas on the GPU: energy,
//                              dE and best state stay consistent, flips count
//                              up, set_beta and copies keep working
// Prints "ok" and exits 0, else prints the failed check and exits 1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "pticm.hh"

using namespace pticm;

static int fails = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("check failed at line %d: %s\n", __LINE__, #c);   \
            ++fails;                                                      \
        }                                                                 \
    } while (0)

static StateVector random_state(Problem& p) {
    std::bernoulli_distribution init_gen(0.5);
    StateVector s(p.vcount());
    for (uint64_t j = 0; j < p.vcount(); ++j) s[j] = init_gen(p.get_rng());
    return s;
}

static int host_checks(const std::string& path) {
    Problem p(path, 7, 0);
    StateVector s = random_state(p);
    std::vector<double> dE(p.vcount());
    double E = p.calculate_dE(dE, s);
    CHECK(E == p.energy(s));
    std::vector<char> raw(p.vcount());
    bool* b = reinterpret_cast<bool*>(raw.data());
    for (uint64_t i = 0; i < p.vcount(); ++i) b[i] = s[i];
    CHECK(p.energy(b) == E);
    std::mt19937_64 g(11);
    for (int k = 0; k < 200; ++k) {  // accept single flips the way Replica::mh_accept does
        const uint32_t i = (uint32_t)(g() % p.vcount());
        E += dE[i];
        p.incremental_dE(dE, s, i);  // before the flip (replica.cc:20)
        s[i] = !s[i];
        std::vector<double> full(p.vcount());
        const double E2 = p.calculate_dE(full, s);
        CHECK(E2 == E);
        CHECK(full == dE);
    }
    for (int k = 0; k < 100; ++k) {
        const double u = p.rand_double();
        CHECK(u >= 0.0 && u < 1.0);
    }
    return fails;
}

static int gpu_checks(const std::string& path) {
    Problem p(path, 3, 0);
    std::vector<double> betas(6);
    beta_linspace(betas.data(), 0.5, 3.0, betas.size(), false);
    std::vector<Replica> replicas(betas.size());
    for (size_t i = 0; i < betas.size(); ++i) {  // initialization.cc:63-76
        StateVector init = random_state(p);
        std::vector<double> dE(p.vcount(), 0);
        const double e = p.calculate_dE(dE, init);
        replicas[i] = Replica(p, betas[i], init, dE, e, true, 5);
    }
    for (int sweep = 0; sweep < 20; ++sweep) {
        for (Replica& r : replicas) {
            const uint64_t f0 = r.get_flips();
            r.sweep(p);
            CHECK(r.get_energy() == p.energy(r.get_state()));
            std::vector<double> dE(p.vcount());
            p.calculate_dE(dE, r.get_state());
            CHECK(dE == r.get_dE());
            CHECK(r.get_flips() >= f0);
            CHECK(r.get_best_ene() == p.energy(r.get_best_state()));
            CHECK(r.get_best_sweep() <= (size_t)sweep);
        }
        if (sweep == 9) {  // a temperature swap the way the reference's exchange applies it
            const double b0 = replicas[0].get_beta();
            replicas[0].set_beta(replicas[1].get_beta());
            replicas[1].set_beta(b0);
        }
    }
    const Replica& cold = replicas.back();
    CHECK(cold.get_flips() > 0);
    CHECK(cold.get_ene_log().size() == 14);  // sweeps 6..19 (sweep_counter > burn_in = 5)
    Replica copy = cold;                      // a copy sweeps on its own
    copy.sweep(p);
    CHECK(copy.get_energy() == p.energy(copy.get_state()));
    return fails;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::printf("usage: facade_caller host|gpu GRAPH\n");
        return 2;
    }
    try {
        const int f = std::strcmp(argv[1], "host") == 0 ? host_checks(argv[2]) : gpu_checks(argv[2]);
        if (f == 0) std::printf("ok\n");
        return f == 0 ? 0 : 1;
    } catch (const Error& e) {
        std::printf("pticm::Error %d: %s\n", e.code, e.what());
        return 1;
    }
}
