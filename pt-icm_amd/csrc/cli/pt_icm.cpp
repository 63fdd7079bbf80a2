// pt_icm.cpp -- the command line of src/main.cc:15-209 on the MI355X engine:
// same flags and defaults (main.cc:33-48), the same echo (main.cc:83-96) and
// the result keys of src/solver.cc:151-163.  Errors print a message and
// return a non-zero status.
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iostream>
#include <string>
#include <vector>

#include "pticm.hh"

namespace {

const char* kUsage =
    "Usage: pt_icm -p PROBLEM [options]\n"
    "  -p, --problem PATH           path to problem file (GSET)            [required]\n"
    "  -s, --seed N                 random seed                            [time(NULL)]\n"
    "  -b, --burn_in N              number of MCMC burn-in steps           [500]\n"
    "      --sweeps N               replica sweeps per PT step             [1]\n"
    "  -r, --replicas N             temperatures per chain                 [20]\n"
    "      --pts N                  parallel tempering steps               [2000]\n"
    "      --preprocess-samples N   pilot samples for --preprocess         [2000]\n"
    "      --preprocess-stepsize X  ladder search step                     [0.1]\n"
    "  -T0, --temp0 X               starting temperature                   [0.6]\n"
    "  -T1, --temp1 X               ending temperature                     [2.8]\n"
    "  -i, --inverse                read T0/T1 as inverse temperatures\n"
    "      --preprocess             multihistogram constant-acceptance ladder\n"
    "      --log                    per-temperature observables T,C,X,u,m after the burn-in\n"
    "                               (written to results.csv, or --log-file)\n"
    "      --icm                    isoenergetic cluster moves\n"
    "      --verbose N              verbosity 0..3                          [0]\n"
    "      --ladder-mode exact|fast|newton  ladder search: reference sums | O(nbins) sums |\n"
    "                             O(nbins) sums + Newton multihistogram solve    [exact]\n"
    "      --log-file PATH          file of --log                          [results.csv]\n"
    "      --ladder-in PATH         inverse temperatures, one per line (instead of -r/-T0/-T1/--preprocess)\n"
    "      --ladder-out PATH        write the inverse temperatures used\n"
    "      --histograms-out PATH    --preprocess: write the pilot's energy histograms (slot,beta,energy,count)\n"
    "      --checkpoint PATH        write the engine state after the run\n"
    "      --resume PATH            continue a run from --checkpoint up to --pts steps in all\n"
    "      --folder DIR             solve every problem file of DIR (batched), --restarts times,\n"
    "      --csv PATH               one CSV row per problem and restart   [results_folder.csv]\n"
    "      --restarts N             restarts of --folder (seed + i)        [1]\n";

struct Args {
    std::string problem, folder, csv = "results_folder.csv";
    uint64_t restarts = 1;
    pticm::Parameters params;
    bool help = false;
};

bool parse(int argc, char** argv, Args& a, std::string& err) {
    pticm::Parameters& p = a.params;
    p.seed = (uint64_t)time(nullptr);
    p.Nt = 20;
    p.pts = 2000;
    p.burn_in = 500;
    p.sweeps = 1;
    p.T0 = 0.6;
    p.T1 = 2.8;
    p.pre_samples = 2000;
    p.pre_step = 0.1;
    for (int i = 1; i < argc; ++i) {
        std::string k = argv[i], v;
        const size_t eq = k.find('=');
        bool has_eq = false;
        if (k.rfind("--", 0) == 0 && eq != std::string::npos) {
            v = k.substr(eq + 1);
            k = k.substr(0, eq);
            has_eq = true;
        }
        auto value = [&](std::string& out) -> bool {
            if (has_eq) { out = v; return true; }
            if (i + 1 >= argc) { err = k + ": missing value"; return false; }
            out = argv[++i];
            return true;
        };
        auto num_u = [&](uint64_t& out) -> bool {
            std::string s;
            if (!value(s)) return false;
            char* end = nullptr;
            out = std::strtoull(s.c_str(), &end, 10);
            if (s.empty() || *end) { err = k + ": expected an unsigned integer, got " + s; return false; }
            return true;
        };
        auto num_f = [&](double& out) -> bool {
            std::string s;
            if (!value(s)) return false;
            char* end = nullptr;
            out = std::strtod(s.c_str(), &end);
            if (s.empty() || *end) { err = k + ": expected a number, got " + s; return false; }
            return true;
        };
        uint64_t u = 0;
        if (k == "-h" || k == "--help") a.help = true;
        else if (k == "-p" || k == "--problem") { if (!value(a.problem)) return false; }
        else if (k == "-s" || k == "--seed") { if (!num_u(p.seed)) return false; }
        else if (k == "-b" || k == "--burn_in") { if (!num_u(p.burn_in)) return false; }
        else if (k == "--sweeps") { if (!num_u(p.sweeps)) return false; }
        else if (k == "-r" || k == "--replicas") { if (!num_u(p.Nt)) return false; }
        else if (k == "--pts") { if (!num_u(p.pts)) return false; }
        else if (k == "--preprocess-samples") { if (!num_u(p.pre_samples)) return false; }
        else if (k == "--preprocess-stepsize") { if (!num_f(p.pre_step)) return false; }
        else if (k == "-T0" || k == "--temp0") { if (!num_f(p.T0)) return false; }
        else if (k == "-T1" || k == "--temp1") { if (!num_f(p.T1)) return false; }
        else if (k == "-i" || k == "--inverse") p.inverse = true;
        else if (k == "--preprocess") p.preprocess = true;
        else if (k == "--log") p.log_data = true;
        else if (k == "--icm") p.icm = true;
        else if (k == "--verbose") { if (!num_u(u)) return false; p.verbosity = (uint32_t)u; }
        else if (k == "--ladder-mode") {
            std::string s;
            if (!value(s)) return false;
            if (s == "exact") p.ladder_mode = PTX_LADDER_EXACT;
            else if (s == "fast") p.ladder_mode = PTX_LADDER_FAST;
            else if (s == "newton") p.ladder_mode = PTX_LADDER_NEWTON;
            else { err = "--ladder-mode: exact, fast or newton"; return false; }
        } else if (k == "--log-file") { if (!value(p.log_path)) return false; }
        else if (k == "--ladder-in") { if (!value(p.ladder_in)) return false; }
        else if (k == "--ladder-out") { if (!value(p.ladder_out)) return false; }
        else if (k == "--histograms-out") { if (!value(p.histograms_out)) return false; }
        else if (k == "--checkpoint") { if (!value(p.checkpoint)) return false; }
        else if (k == "--resume") { if (!value(p.resume)) return false; }
        else if (k == "--folder") { if (!value(a.folder)) return false; }
        else if (k == "--csv") { if (!value(a.csv)) return false; }
        else if (k == "--restarts") { if (!num_u(a.restarts)) return false; }
        else { err = "Unknown argument: " + k; return false; }
    }
    if (!a.help && a.problem.empty() && a.folder.empty()) { err = "-p: required."; return false; }
    if (!a.help && p.Nt < 1) { err = "-r: at least one temperature"; return false; }
    if (!a.help && p.preprocess && p.Nt < 2) { err = "--preprocess needs at least 2 temperatures"; return false; }
    return true;
}

void print_args(const pticm::Parameters& p) {  // main.cc:83-96
    std::cout << "Seed: " << p.seed << std::endl;
    std::cout << "Replicas: " << p.Nt << std::endl;
    std::cout << "Sweeps: " << p.sweeps << std::endl;
    std::cout << "PT Steps: " << p.pts << std::endl;
    std::cout << "Burn-in: " << p.burn_in << std::endl;
    std::cout << "T0: " << p.T0 << std::endl;
    std::cout << "T1: " << p.T1 << std::endl;
    std::cout << "Inverse: " << p.inverse << std::endl;
    std::cout << "Preprocess: " << p.preprocess << std::endl;
    std::cout << "Log Data: " << p.log_data << std::endl;
    std::cout << "ICM: " << p.icm << std::endl;
    std::cout << "Verbosity: " << p.verbosity << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    std::string err;
    if (!parse(argc, argv, a, err)) {
        std::cout << err << std::endl << kUsage;
        return 1;
    }
    if (a.help) {
        std::cout << kUsage;
        return 0;
    }
    print_args(a.params);
    std::cout << a.problem << std::endl;
    try {
        pticm::PTWorker w;
        w.params = a.params;
        if (!a.folder.empty()) {
            pticm::solve_folder(w, a.folder, a.csv, a.restarts);
            return 0;
        }
        double best_ene = 0;
        size_t flips = 0;
        pticm::StateVector best_state;
        pticm::solve(w, a.problem, best_ene, flips, best_state);
    } catch (const pticm::Error& e) {
        std::fprintf(stderr, "pt_icm: error %d: %s\n", e.code, e.what());
        return 2;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "pt_icm: %s\n", e.what());
        return 2;
    }
    return 0;
}
