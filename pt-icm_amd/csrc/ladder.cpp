// ladder.cpp -- see ladder.hpp.  All arithmetic is FP64 on the host; the
// exact mode follows the reference's operation order so that the same
// histograms give the same ladder bit for bit.
#include "ladder.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "ptx_common.h"

namespace ptx {

void beta_linspace(double* temps, double T0, double T1, int64_t n, bool inverse) {
    double step, curr;
    if (inverse) {
        step = T1 < T0 ? (T0 - T1) / (double)(n - 1) : (T1 - T0) / (double)(n - 1);
        curr = T1 < T0 ? T1 : T0;
    } else {
        step = T1 < T0 ? (1 / T1 - 1 / T0) / (double)(n - 1) : (1 / T0 - 1 / T1) / (double)(n - 1);
        curr = T1 < T0 ? 1 / T0 : 1 / T1;
    }
    for (int64_t i = 0; i < n; ++i) {
        temps[i] = curr;
        curr += step;
    }
}

int32_t icm_count(const double* beta, int32_t nt, double char_scale) {
    int32_t nc = 0;
    for (int32_t i = nt - 1; i >= 0 && !(beta[i] < char_scale); --i) ++nc;
    return nc;
}

void pilot_betas(double T0, double T1, int32_t nt, bool inverse, double* out, double* bs_out, double* be_out) {
    double bs = inverse ? T0 : 1 / T0;
    double be = inverse ? T1 : 1 / T1;
    if (bs < be) std::swap(bs, be);
    const double step = (be - bs) / (double)(nt - 1);
    double b = bs;
    for (int32_t i = 0; i < nt; ++i, b += step) out[i] = b;
    *bs_out = bs;
    *be_out = be;
}

void threshold_table(const double* beta, int32_t nt, int32_t fmax, uint32_t* thr) {
    for (int32_t t = 0; t < nt; ++t)
        for (int32_t m = 1; m <= fmax; ++m) {
            const double p = pexp(-(beta[t] * (double)(2 * m))) * 4294967296.0;
            const double fl = std::floor(p);
            thr[(size_t)t * fmax + (m - 1)] = fl >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)fl;
        }
}

namespace {

inline double log_add_count(double count, double l2) {  // logplusexp, cae_estimator.cc:19-26
    double l1 = std::log(count);
    if (l1 < l2) std::swap(l1, l2);
    return l1 + std::log1p(std::exp(l2 - l1));
}
inline double log_add(double l1, double l2) {  // logplusexp2, cae_estimator.cc:28-35
    if (l1 < l2) std::swap(l1, l2);
    return l1 + std::log1p(std::exp(l2 - l1));
}
inline size_t bin_of(double e, double emin, double estep) { return (size_t)std::floor((e - emin) / estep); }

void pe_vec(const std::vector<double>& logdos, std::vector<double>& out, double emin, double beta, double logz,
            double estep) {
    out.resize(logdos.size());
    double e = emin;
    for (size_t i = 0; i < logdos.size(); ++i, e += estep) out[i] = logdos[i] - e * beta - logz;
}

}  // namespace

std::vector<SlotHist> rebin(const uint32_t* counts, int32_t nt, int64_t nbins, int64_t e_base, double estep) {
    std::vector<SlotHist> out(nt);
    for (int32_t s = 0; s < nt; ++s) {
        const uint32_t* c = counts + (size_t)s * nbins;
        int64_t lo = -1, hi = -1;
        for (int64_t b = 0; b < nbins; ++b)
            if (c[b]) {
                if (lo < 0) lo = b;
                hi = b;
            }
        SlotHist& H = out[s];
        if (lo < 0) continue;
        H.emin = (double)(e_base + 2 * lo);
        H.emax = (double)(e_base + 2 * hi);
        const size_t nb = (size_t)std::ceil((H.emax - H.emin) / estep) + 1;
        H.h.assign(nb, 0);
        for (int64_t b = lo; b <= hi; ++b)
            if (c[b]) H.h[bin_of((double)(e_base + 2 * b), H.emin, estep)] += c[b];
    }
    return out;
}

double partition_func(const std::vector<double>& logdos, double beta, double emin, double estep) {
    double logz = 1;
    for (size_t i = 0; i < logdos.size(); ++i) {
        const double e = (double)i * estep + emin;
        const double lb = -beta * e;
        logz = (i == 0) ? lb + logdos[i] : log_add(logz, lb + logdos[i]);
    }
    return logz;
}

Dos solve_dos(const std::vector<SlotHist>& hist, const double* beta_h, double estep) {
    Dos D;
    D.estep = estep;
    const size_t nh = hist.size();
    D.max_emax = -DBL_MAX;
    D.min_emin = DBL_MAX;
    for (const auto& H : hist) {
        D.max_emax = std::max(D.max_emax, H.emax);
        D.min_emin = std::min(D.min_emin, H.emin);
    }
    uint64_t n = 0;  // entries per histogram (the reference takes histogram 0's total)
    for (uint32_t c : hist[0].h) n += c;
    const size_t nbins = (size_t)std::ceil((D.max_emax - D.min_emin) / estep);  // drops the top bin on exact spans
    D.logdos.assign(nbins, 0.0);
    D.logz.assign(nh, 0.0);
    const double logn = std::log((double)n);
    double max_diff;
    do {
        max_diff = 0;
        double e = D.min_emin;
        for (size_t i = 0; i < nbins; ++i, e += estep) {
            double lognum = 1, logdenom = 1;
            bool firstnum = true;
            for (size_t j = 0; j < nh; ++j) {
                const double term = logn - D.logz[j] - beta_h[j] * e;
                logdenom = (j == 0) ? term : log_add(term, logdenom);
                if (e < hist[j].emin) continue;
                const double idx = (double)bin_of(e, hist[j].emin, estep);
                if (idx >= (double)hist[j].h.size()) continue;
                const uint32_t c = hist[j].h[(size_t)idx];
                if (!c) continue;
                if (firstnum) {
                    firstnum = false;
                    lognum = std::log((double)c);
                } else {
                    lognum = log_add_count((double)c, lognum);
                }
            }
            D.logdos[i] = firstnum ? -DBL_MAX : lognum - logdenom;
        }
        for (size_t j = 0; j < nh; ++j) {
            const double tz = partition_func(D.logdos, beta_h[j], D.min_emin, estep);
            max_diff += std::pow(std::fabs(D.logz[j] - tz) / std::fabs(tz), 2);
            D.logz[j] = tz;
        }
    } while (max_diff > 1e-7);
    return D;
}

// The same fixed point as solve_dos, found by damped Newton on the convex
// function whose stationarity conditions are the multihistogram equations:
//   F(g) = sum_e N(e) log sum_j n e^{g_j - beta_j e} - n sum_j g_j,
// g_j = -log Z_j (gauge g_0 = 0), N(e) = all counts in the energy bin.
// Gradient_k = sum_e N(e) p_k(e) - n, Hessian_kl = sum_e N(e)(p_k d_kl - p_k p_l)
// with p_k(e) the normalised weights e^{g_k - beta_k e} / sum_j.  Starts from
// thermodynamic integration of the slot means (d log Z / d beta = -<E>).
Dos solve_dos_newton(const std::vector<SlotHist>& hist, const double* beta_h, double estep) {
    Dos D;
    D.estep = estep;
    const int nh = (int)hist.size();
    D.max_emax = -DBL_MAX;
    D.min_emin = DBL_MAX;
    for (const auto& H : hist) {
        D.max_emax = std::max(D.max_emax, H.emax);
        D.min_emin = std::min(D.min_emin, H.emin);
    }
    uint64_t n = 0;
    for (uint32_t c : hist[0].h) n += c;
    const size_t nbins = (size_t)std::ceil((D.max_emax - D.min_emin) / estep);
    std::vector<double> N(nbins, 0.0);
    std::vector<double> mean(nh, 0.0);
    for (int j = 0; j < nh; ++j) {
        const SlotHist& H = hist[j];
        double s = 0, w = 0;
        for (size_t b = 0; b < H.h.size(); ++b) {
            if (!H.h[b]) continue;
            const double e = H.emin + (double)b * estep;
            s += e * H.h[b];
            w += H.h[b];
            const size_t i = bin_of(e, D.min_emin, estep);
            if (i < nbins) N[i] += H.h[b];
        }
        mean[j] = w > 0 ? s / w : 0.0;
    }
    std::vector<double> E, Nv;  // occupied bins only
    for (size_t i = 0; i < nbins; ++i)
        if (N[i] > 0) {
            E.push_back(D.min_emin + (double)i * estep);
            Nv.push_back(N[i]);
        }
    const double ln_n = std::log((double)n), dn = (double)n;
    std::vector<double> g(nh, 0.0);
    for (int j = 1; j < nh; ++j) g[j] = g[j - 1] + (beta_h[j] - beta_h[j - 1]) * 0.5 * (mean[j] + mean[j - 1]);
    std::vector<double> t(nh), p(nh), grad(nh), Hm((size_t)nh * nh), step(nh), gt(nh);
    auto objective = [&](const std::vector<double>& gg) {
        double F = 0;
        for (size_t i = 0; i < E.size(); ++i) {
            double mx = -DBL_MAX;
            for (int j = 0; j < nh; ++j) mx = std::max(mx, gg[j] - beta_h[j] * E[i]);
            double s = 0;
            for (int j = 0; j < nh; ++j) s += std::exp(gg[j] - beta_h[j] * E[i] - mx);
            F += Nv[i] * (ln_n + mx + std::log(s));
        }
        for (int j = 0; j < nh; ++j) F -= dn * gg[j];
        return F;
    };
    double F = objective(g);
    for (int iter = 0; iter < 200; ++iter) {
        std::fill(grad.begin(), grad.end(), -dn);
        std::fill(Hm.begin(), Hm.end(), 0.0);
        std::vector<int> live;
        live.reserve(nh);
        for (size_t i = 0; i < E.size(); ++i) {
            double mx = -DBL_MAX;
            for (int j = 0; j < nh; ++j) mx = std::max(mx, t[j] = g[j] - beta_h[j] * E[i]);
            double s = 0;
            for (int j = 0; j < nh; ++j) s += (p[j] = std::exp(t[j] - mx));
            const double inv = 1.0 / s;
            live.clear();
            for (int j = 0; j < nh; ++j) {
                p[j] *= inv;
                if (p[j] > 1e-17) live.push_back(j);  // a bin weighs on a few temperatures only
            }
            for (int k = 0; k < nh; ++k) grad[k] += Nv[i] * p[k];
            for (int k : live) {
                const double a = Nv[i] * p[k];
                Hm[(size_t)k * nh + k] += a;
                double* row = &Hm[(size_t)k * nh];
                for (int l : live) row[l] -= a * p[l];
            }
        }
        double gmax = 0;
        for (int k = 1; k < nh; ++k) gmax = std::max(gmax, std::fabs(grad[k]));
        if (gmax <= 1e-10 * dn) break;
        // solve H' step = -grad' on slots 1..nh-1 (Cholesky with a tiny ridge)
        const int m = nh - 1;
        std::vector<double> A((size_t)m * m), b(m);
        double tr = 0;
        for (int k = 0; k < m; ++k) tr += Hm[(size_t)(k + 1) * nh + (k + 1)];
        const double ridge = 1e-12 * (tr > 0 ? tr / m : 1.0);
        for (int k = 0; k < m; ++k) {
            b[k] = -grad[k + 1];
            for (int l = 0; l < m; ++l) A[(size_t)k * m + l] = Hm[(size_t)(k + 1) * nh + (l + 1)];
            A[(size_t)k * m + k] += ridge;
        }
        bool ok = true;
        for (int k = 0; k < m && ok; ++k) {
            for (int l = 0; l < k; ++l) A[(size_t)k * m + k] -= A[(size_t)k * m + l] * A[(size_t)k * m + l];
            if (!(A[(size_t)k * m + k] > 0)) {
                ok = false;
                break;
            }
            A[(size_t)k * m + k] = std::sqrt(A[(size_t)k * m + k]);
            for (int r = k + 1; r < m; ++r) {
                double v = A[(size_t)r * m + k];
                for (int l = 0; l < k; ++l) v -= A[(size_t)r * m + l] * A[(size_t)k * m + l];
                A[(size_t)r * m + k] = v / A[(size_t)k * m + k];
            }
        }
        if (!ok) break;
        for (int k = 0; k < m; ++k) {  // forward then backward substitution
            double v = b[k];
            for (int l = 0; l < k; ++l) v -= A[(size_t)k * m + l] * b[l];
            b[k] = v / A[(size_t)k * m + k];
        }
        for (int k = m - 1; k >= 0; --k) {
            double v = b[k];
            for (int l = k + 1; l < m; ++l) v -= A[(size_t)l * m + k] * b[l];
            b[k] = v / A[(size_t)k * m + k];
        }
        step[0] = 0;
        for (int k = 0; k < m; ++k) step[k + 1] = b[k];
        double a = 1.0, Fn = F;
        bool moved = false;
        for (int ls = 0; ls < 40; ++ls, a *= 0.5) {
            for (int j = 0; j < nh; ++j) gt[j] = g[j] + a * step[j];
            Fn = objective(gt);
            if (Fn <= F) {
                moved = true;
                break;
            }
        }
        if (!moved) break;
        const double dF = F - Fn;
        g.swap(gt);
        F = Fn;
        if (dF <= 1e-15 * std::fabs(F)) break;
    }
    // log g(e) and log Z_j as the reference computes them from the fixed point
    D.logdos.assign(nbins, -DBL_MAX);
    for (size_t i = 0; i < nbins; ++i) {
        if (!(N[i] > 0)) continue;
        const double e = D.min_emin + (double)i * estep;
        double mx = -DBL_MAX;
        for (int j = 0; j < nh; ++j) mx = std::max(mx, g[j] - beta_h[j] * e);
        double s = 0;
        for (int j = 0; j < nh; ++j) s += std::exp(g[j] - beta_h[j] * e - mx);
        D.logdos[i] = std::log(N[i]) - (ln_n + mx + std::log(s));
    }
    D.logz.resize(nh);
    for (int j = 0; j < nh; ++j) D.logz[j] = partition_func(D.logdos, beta_h[j], D.min_emin, estep);
    return D;
}

double acceptance_rate(const std::vector<double>& p1, const std::vector<double>& p2, double beta_1, double beta_2,
                       double estep, double emin, bool fast) {
    const size_t nb = p1.size();
    if (!fast) {
        // the reference's double sum, same association order
        double r = 1;
        bool first = true;
        double e1 = emin;
        for (size_t i = 0; i < nb; ++i, e1 += estep) {
            const double p1i = p1[i];
            for (size_t j = 0; j <= i; ++j) {
                r = first ? p1i + p2[j] : log_add(r, p1i + p2[j]);
                first = false;
            }
            double e2 = e1 + estep;
            for (size_t j = i + 1; j < nb; ++j, e2 += estep) {
                const double p12 = (beta_1 - beta_2) * (e1 - e2);
                r = first ? p1i + p2[j] + p12 : log_add(r, p1i + p2[j] + p12);
                first = false;
            }
        }
        return std::exp(r);
    }
    // r = sum_i P1_i [ sum_{j<=i} P2_j + e^{db e_i} sum_{j>i} P2_j e^{-db e_j} ], db = beta_1 - beta_2,
    // with running log-sum-exps (prefix for the first sum, suffix for the second).
    const double db = beta_1 - beta_2;
    const double skip = -DBL_MAX / 4;  // empty bins carry -DBL_MAX and contribute nothing
    std::vector<double> suffix(nb + 1, -INFINITY);
    double e = emin + (double)(nb - 1) * estep;
    for (size_t k = nb; k-- > 0; e -= estep) {
        const double v = p2[k] < skip ? -INFINITY : p2[k] - db * e;
        if (v == -INFINITY) suffix[k] = suffix[k + 1];
        else suffix[k] = suffix[k + 1] == -INFINITY ? v : log_add(suffix[k + 1], v);
    }
    double prefix = -INFINITY, total = -INFINITY;
    e = emin;
    for (size_t i = 0; i < nb; ++i, e += estep) {
        if (!(p2[i] < skip)) prefix = prefix == -INFINITY ? p2[i] : log_add(prefix, p2[i]);
        if (p1[i] < skip) continue;
        double inner = prefix;
        if (suffix[i + 1] != -INFINITY) {
            const double up = db * e + suffix[i + 1];
            inner = inner == -INFINITY ? up : log_add(inner, up);
        }
        if (inner == -INFINITY) continue;
        const double term = p1[i] + inner;
        total = total == -INFINITY ? term : log_add(total, term);
    }
    return std::exp(total);
}

int32_t ladder_search(const Dos& D, int32_t nt, double beta_start, double beta_end, double step_size, double r,
                      bool fast, double* bv) {
    const double estep = D.estep, emin = D.min_emin;
    bv[0] = beta_start;
    double part_last = partition_func(D.logdos, beta_start, emin, estep);
    double step = step_size;
    std::vector<double> prev, curr;
    pe_vec(D.logdos, prev, emin, beta_start, part_last, estep);
    int32_t i;
    for (i = 1; i < nt; ++i) {
        const double beta_prev = bv[i - 1];
        double beta_curr = beta_prev;
        double part_curr = 0, r_curr = 0, r_last = -1;
        bool cont;
        do {
            cont = false;
            const double beta_orig = beta_curr;
            beta_curr = 1 / (1 / beta_curr + step);
            part_curr = partition_func(D.logdos, beta_curr, emin, estep);
            pe_vec(D.logdos, curr, emin, beta_curr, part_curr, estep);
            r_curr = acceptance_rate(prev, curr, beta_prev, beta_curr, estep, emin, fast);
            if ((r - r_curr) > 0.05) {  // overshoot: back off and halve the step
                beta_curr = beta_orig;
                step /= 2;
                cont = true;
                continue;
            }
            if (r_last > 0 && std::fabs(r_curr - r_last) < 5e-2) {  // plateau: double the step
                step *= 2;
                continue;
            }
            r_last = r_curr;
        } while ((r_curr > r && beta_curr > beta_end) || cont);
        bv[i] = beta_curr;
        if (beta_curr <= beta_end) break;
        part_last = part_curr;
        std::swap(prev, curr);
    }
    return i;
}

}  // namespace ptx
