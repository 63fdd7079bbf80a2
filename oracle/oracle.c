/*
 * oracle.c -- CPU restatement of the PT-ICM hot path.  TEST INFRASTRUCTURE:
 * loaded only by tests/, the smoke check and bench.py's CPU-baseline leg, as
 * the checker; never by the product.  Built with -ffp-contract=off so every
 * FP64 expression rounds exactly as written.
 *
 * How each part is pinned (tests/test_oracle*.py):
 *   or_philox4x32_10   Random123 known-answer vectors (Salmon et al. 2011).
 *   or_exp / or_log    against libm over the full ranges used (rel err < 4 ulp)
 *                      and exact special values (exp(0)=1, log(1)=0, powers of 2).
 *   or_read_gset       closed forms on the reference's example graphs
 *                      (lattice all-up energy -2N, M/degree counts), the
 *                      reference's published G022 state (E=-6720, cut=13,355,
 *                      scripts/grab_results2csv.py:24-28), 0/1-based pairs.
 *   or_energy/maxcut   brute force on tiny graphs, cut = (sum w - E)/2.
 *   or_delta_e, or_incremental_de  the invariant of src/problem.cc:64-100
 *                      (dE_i = E(flip i) - E for every i and every i,j pair).
 *   or_colour          proper colouring (no intra-class edge), 2 classes on
 *                      even tori and bipartite graphs (DSatur is exact there).
 *   or_beta_linspace   the reference's printed ladder in
 *                      scripts/grab_results2csv.py:31.
 *   engine sweep       exact Boltzmann distribution by enumeration (chi2) on
 *                      N<=10 instances in both sweep orders; 2D ferromagnet u(T)
 *                      against data/lattice_2d_20_u.csv.
 *   engine exchange    permutation invariant, per-slot rates against the
 *                      closed form min(1, e^{dB dE}) averaged over exact DOS.
 *   engine ICM         E1+E2 conserved (h=0), cluster = union-find component,
 *                      spins outside the cluster unchanged, global-flip rule.
 *   DOS / ladder       exact DOS of the 4x4 / 5x5 torus by enumeration; WHAM
 *                      recovers it; acceptance_rate against a direct double sum.
 */
#include "oracle.h"

#include <ctype.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw 2011, "Parallel random numbers:  */
/* as easy as 1, 2, 3").                                                      */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream tags (low byte of counter word 3; the instance id fills the rest). */
enum { TAG_INIT = 1, TAG_SWEEP = 2, TAG_EXCH = 3, TAG_SHUF = 4, TAG_ICM = 5, TAG_REFSWEEP = 6,
       TAG_PILOT = 0x10 };

static uint32_t draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t tag,
                     uint32_t inst, uint32_t word) {
    uint32_t ctr[4] = {c0, c1, c2, (tag & 0xFFu) | (inst << 8)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    or_philox4x32_10(ctr, key, out);
    return out[word & 3u];
}

/* ------------------------------------------------------------------------ */
/* Portable FP64 exp/log: only +, -, *, floor and exponent bit manipulation, */
/* so that any IEEE-754 binary64 implementation without contraction gives    */
/* the same bits.                                                            */
/* ------------------------------------------------------------------------ */
static double bits_to_double(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static uint64_t double_to_bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* exp: Cody & Waite (1980) argument reduction x = k ln2 + r with the fdlibm
 * two-part ln2, then the degree-14 Taylor polynomial of e^r, |r| <= 0.35. */
double or_exp(double x) {
    if (x != x) return x;
    if (x > 709.0) return INFINITY;
    if (x < -745.0) return 0.0;
    const double inv_ln2 = 1.44269504088896338700e+00;
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    double kd = floor(x * inv_ln2 + 0.5);
    double r = (x - kd * ln2_hi) - kd * ln2_lo;
    static const double inv_fact[15] = {
        1.0, 1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0, 1.0 / 5040.0,
        1.0 / 40320.0, 1.0 / 362880.0, 1.0 / 3628800.0, 1.0 / 39916800.0, 1.0 / 479001600.0,
        1.0 / 6227020800.0, 1.0 / 87178291200.0};
    double p = inv_fact[14];
    for (int i = 13; i >= 0; --i) p = p * r + inv_fact[i];
    int k = (int)kd;
    if (k < -1022) return 0.0; /* results below the normal range are flushed */
    if (k > 1023) return INFINITY;
    return p * bits_to_double((uint64_t)(k + 1023) << 52);
}

static const double k_pow2_neg_j32[32] = {
    0x1.0000000000000p+0, 0x1.f50765b6e4540p-1, 0x1.ea4afa2a490dap-1, 0x1.dfc97337b9b5fp-1,
    0x1.d5818dcfba487p-1, 0x1.cb720dcef9069p-1, 0x1.c199bdd85529cp-1, 0x1.b7f76f2fb5e47p-1,
    0x1.ae89f995ad3adp-1, 0x1.a5503b23e255dp-1, 0x1.9c49182a3f090p-1, 0x1.93737b0cdc5e5p-1,
    0x1.8ace5422aa0dbp-1, 0x1.82589994cce13p-1, 0x1.7a11473eb0187p-1, 0x1.71f75e8ec5f74p-1,
    0x1.6a09e667f3bcdp-1, 0x1.6247eb03a5585p-1, 0x1.5ab07dd485429p-1, 0x1.5342b569d4f82p-1,
    0x1.4bfdad5362a27p-1, 0x1.44e086061892dp-1, 0x1.3dea64c123422p-1, 0x1.371a7373aa9cbp-1,
    0x1.306fe0a31b715p-1, 0x1.29e9df51fdee1p-1, 0x1.2387a6e756238p-1, 0x1.1d4873168b9aap-1,
    0x1.172b83c7d517bp-1, 0x1.11301d0125b51p-1, 0x1.0b5586cf9890fp-1, 0x1.059b0d3158574p-1};

/* log: x = 2^k m, m in [1,2); j = the largest index with m c_j >= 1 where
 * c_j = 2^(-j/32) (binary search over the table); y = m c_j in
 * [1, 2^(1/32)); log x = (k + j/32) ln2 + log1p(y-1) with the degree-13
 * Taylor series of log1p (0 <= y-1 < 0.0222).  0 -> -inf, negative -> NaN,
 * subnormals are scaled first. */
double or_log(double x) {
    if (x != x || x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    int k = 0;
    uint64_t b = double_to_bits(x);
    if ((b >> 52) == 0) { /* subnormal */
        x = x * 0x1p54;
        k = -54;
        b = double_to_bits(x);
    }
    k += (int)((b >> 52) & 0x7FF) - 1023;
    double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
    int lo = 0, hi = 31;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (m * k_pow2_neg_j32[mid] >= 1.0) lo = mid;
        else hi = mid - 1;
    }
    int j = lo;
    double t = m * k_pow2_neg_j32[j] - 1.0;
    static const double c[14] = {0.0,        1.0,        -1.0 / 2.0,  1.0 / 3.0,  -1.0 / 4.0,
                                 1.0 / 5.0,  -1.0 / 6.0, 1.0 / 7.0,   -1.0 / 8.0, 1.0 / 9.0,
                                 -1.0 / 10.0, 1.0 / 11.0, -1.0 / 12.0, 1.0 / 13.0};
    double p = c[13];
    for (int i = 12; i >= 1; --i) p = p * t + c[i];
    p = p * t;
    const double ln2 = 6.93147180559945286227e-01;
    double kj = (double)k + (double)j * 0.03125;
    return kj * ln2 + p;
}

double or_uniform_log(uint32_t r) {
    if (r == 0) return -INFINITY;
    return or_log((double)r * 0x1p-32);
}

/* ------------------------------------------------------------------------ */
/* Problem                                                                   */
/* ------------------------------------------------------------------------ */
struct or_graph {
    int32_t n;
    int64_t m_header;
    int sparse;
    double char_scale;
    int32_t* rowptr; /* n+1 */
    int32_t* col;    /* nnz */
    int32_t* w;      /* nnz */
    int32_t* h;      /* n */
};

static void set_err(char* err, int len, const char* msg) {
    if (err && len > 0) {
        strncpy(err, msg, (size_t)len - 1);
        err[len - 1] = 0;
    }
}

static int is_integral(double x) { return x == floor(x) && fabs(x) < 1e9; }

/* Builds the CSR from an edge list in file order.  Sparse: every listed entry
 * is kept (duplicates and zero weights included), each stored at both ends in
 * list order (problem.cc:226-229).  Dense: matrix semantics, the last weight
 * of a (u,v) pair wins, zero entries are no edge (problem.cc:230-232, and the
 * |w|>1e-10 test in icm.cc:62). */
static or_graph* build_graph(int32_t n, int64_t m_header, int sparse, int64_t cnt, const int64_t* eu,
                             const int64_t* ev, const double* ew, const double* h, double char_scale,
                             char* err, int errlen) {
    or_graph* g = (or_graph*)calloc(1, sizeof(or_graph));
    g->n = n;
    g->m_header = m_header;
    g->sparse = sparse;
    g->char_scale = char_scale;
    g->h = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    for (int32_t i = 0; i < n; ++i) {
        if (!is_integral(h[i])) { set_err(err, errlen, "non-integral linear term"); or_graph_free(g); return NULL; }
        g->h[i] = (int32_t)h[i];
    }
    for (int64_t e = 0; e < cnt; ++e) {
        if (eu[e] < 0 || eu[e] >= n || ev[e] < 0 || ev[e] >= n) {
            set_err(err, errlen, "edge endpoint out of range"); or_graph_free(g); return NULL;
        }
        if (!is_integral(ew[e])) { set_err(err, errlen, "non-integral edge weight"); or_graph_free(g); return NULL; }
    }
    int64_t* deg = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    if (sparse) {
        for (int64_t e = 0; e < cnt; ++e) {
            if (eu[e] == ev[e]) { free(deg); set_err(err, errlen, "self-loop in a sparse instance"); or_graph_free(g); return NULL; }
            deg[eu[e]]++; deg[ev[e]]++;
        }
        g->rowptr = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
        g->rowptr[0] = 0;
        for (int32_t i = 0; i < n; ++i) g->rowptr[i + 1] = g->rowptr[i] + (int32_t)deg[i];
        int64_t nnz = g->rowptr[n];
        g->col = (int32_t*)malloc((size_t)(nnz + 1) * sizeof(int32_t));
        g->w = (int32_t*)malloc((size_t)(nnz + 1) * sizeof(int32_t));
        int64_t* fill = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
        for (int64_t e = 0; e < cnt; ++e) {
            int64_t u = eu[e], v = ev[e];
            int64_t pu = g->rowptr[u] + fill[u]++;
            g->col[pu] = (int32_t)v; g->w[pu] = (int32_t)ew[e];
            int64_t pv = g->rowptr[v] + fill[v]++;
            g->col[pv] = (int32_t)u; g->w[pv] = (int32_t)ew[e];
        }
        free(fill);
    } else {
        double* mat = (double*)calloc((size_t)n * (size_t)n, sizeof(double));
        for (int64_t e = 0; e < cnt; ++e) {
            mat[(size_t)eu[e] * n + ev[e]] = ew[e];
            mat[(size_t)ev[e] * n + eu[e]] = ew[e];
        }
        for (int32_t i = 0; i < n; ++i)
            for (int32_t j = 0; j < n; ++j)
                if (i != j && fabs(mat[(size_t)i * n + j]) > 1e-10) deg[i]++;
        g->rowptr = (int32_t*)malloc(((size_t)n + 1) * sizeof(int32_t));
        g->rowptr[0] = 0;
        for (int32_t i = 0; i < n; ++i) g->rowptr[i + 1] = g->rowptr[i] + (int32_t)deg[i];
        int64_t nnz = g->rowptr[n];
        g->col = (int32_t*)malloc((size_t)(nnz + 1) * sizeof(int32_t));
        g->w = (int32_t*)malloc((size_t)(nnz + 1) * sizeof(int32_t));
        int64_t p = 0;
        for (int32_t i = 0; i < n; ++i)
            for (int32_t j = 0; j < n; ++j)
                if (i != j && fabs(mat[(size_t)i * n + j]) > 1e-10) {
                    g->col[p] = j; g->w[p] = (int32_t)mat[(size_t)i * n + j]; ++p;
                }
        free(mat);
    }
    free(deg);
    return g;
}

or_graph* or_read_gset(const char* path, char* err, int errlen) {
    FILE* f = fopen(path, "r");
    if (!f) { set_err(err, errlen, "cannot open problem file"); return NULL; }
    int is_ising = strstr(path, ".ising") != NULL;
    int zero_based = strstr(path, ".gset") != NULL || is_ising;
    long long N = 0, L = 0, M = 0;
    int got = is_ising ? fscanf(f, "%lld %lld %lld", &N, &L, &M) : fscanf(f, "%lld %lld", &N, &M);
    if (got != (is_ising ? 3 : 2) || N <= 0 || M < 0 || L < 0) {
        fclose(f); set_err(err, errlen, "malformed header"); return NULL;
    }
    /* problem.cc:198 */
    int sparse = ((double)M / (double)(N * (N - 1) / 2)) < 0.1;
    double* h = (double*)calloc((size_t)N, sizeof(double));
    int64_t cap = M > 0 ? M : 1, cnt = 0;
    int64_t* eu = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
    int64_t* ev = (int64_t*)malloc((size_t)cap * sizeof(int64_t));
    double* ew = (double*)malloc((size_t)cap * sizeof(double));
    double sum = 0.0, sqsum = 0.0;
    long long nlin = 0;
    char line[4096];
    int off = zero_based ? 0 : -1;
    /* rest of the header line */
    if (!fgets(line, sizeof line, f)) line[0] = 0;
    while (fgets(line, sizeof line, f)) {
        size_t len = strlen(line);
        if (len > 0 && line[len - 1] == '\n') line[--len] = 0;
        if (len <= 1) continue; /* problem.cc:211 skips lines of size <= 1 */
        char* s = line;
        char* end;
        if (nlin < L) {
            long long u = strtoll(s, &end, 10);
            if (end == s) goto malformed;
            s = end;
            double w = strtod(s, &end);
            if (end == s) goto malformed;
            u += off;
            if (u < 0 || u >= N) goto malformed;
            h[u] = w;
            ++nlin;
            continue;
        }
        long long u = strtoll(s, &end, 10);
        if (end == s) goto malformed;
        s = end;
        long long v = strtoll(s, &end, 10);
        if (end == s) goto malformed;
        s = end;
        double w = strtod(s, &end);
        if (end == s) w = 1.0; /* problem.cc:229-233: missing weight means 1 */
        u += off; v += off;
        if (u < 0 || u >= N || v < 0 || v >= N) goto malformed;
        if (cnt == cap) {
            cap *= 2;
            eu = (int64_t*)realloc(eu, (size_t)cap * sizeof(int64_t));
            ev = (int64_t*)realloc(ev, (size_t)cap * sizeof(int64_t));
            ew = (double*)realloc(ew, (size_t)cap * sizeof(double));
        }
        eu[cnt] = u; ev[cnt] = v; ew[cnt] = w; ++cnt;
        sum += w;
        sqsum += w * w;
    }
    fclose(f);
    {
        /* problem.cc:244-245 (uses M from the header) */
        double variance = sqsum / M - (sum / M * sum / M);
        double mean = sum / M;
        double cs = mean < variance ? variance : mean; /* std::max */
        or_graph* g = build_graph((int32_t)N, M, sparse, cnt, eu, ev, ew, h, cs, err, errlen);
        free(eu); free(ev); free(ew); free(h);
        return g;
    }
malformed:
    fclose(f);
    free(eu); free(ev); free(ew); free(h);
    set_err(err, errlen, "malformed edge or linear-term line");
    return NULL;
}

or_graph* or_graph_from_coo(int32_t n, int64_t m, const int32_t* u, const int32_t* v, const double* w,
                            const double* h, int dense, char* err, int errlen) {
    int64_t* eu = (int64_t*)malloc((size_t)(m + 1) * sizeof(int64_t));
    int64_t* ev = (int64_t*)malloc((size_t)(m + 1) * sizeof(int64_t));
    double* hh = (double*)calloc((size_t)n + 1, sizeof(double));
    double sum = 0, sq = 0;
    for (int64_t e = 0; e < m; ++e) { eu[e] = u[e]; ev[e] = v[e]; sum += w[e]; sq += w[e] * w[e]; }
    if (h) for (int32_t i = 0; i < n; ++i) hh[i] = h[i];
    double var = m > 0 ? sq / m - (sum / m * sum / m) : 0.0;
    double cs = m > 0 ? fmax(sum / m, var) : 0.0;
    or_graph* g = build_graph(n, m, !dense, m, eu, ev, w, hh, cs, err, errlen);
    free(eu); free(ev); free(hh);
    return g;
}

void or_graph_free(or_graph* g) {
    if (!g) return;
    free(g->rowptr); free(g->col); free(g->w); free(g->h); free(g);
}
int32_t or_graph_n(const or_graph* g) { return g->n; }
int64_t or_graph_nnz(const or_graph* g) { return g->rowptr[g->n]; }
int64_t or_graph_m_header(const or_graph* g) { return g->m_header; }
int or_graph_sparse(const or_graph* g) { return g->sparse; }
double or_graph_char_scale(const or_graph* g) { return g->char_scale; }
void or_graph_csr(const or_graph* g, int32_t* rowptr, int32_t* col, int32_t* w, int32_t* h) {
    memcpy(rowptr, g->rowptr, ((size_t)g->n + 1) * sizeof(int32_t));
    memcpy(col, g->col, (size_t)g->rowptr[g->n] * sizeof(int32_t));
    memcpy(w, g->w, (size_t)g->rowptr[g->n] * sizeof(int32_t));
    memcpy(h, g->h, (size_t)g->n * sizeof(int32_t));
}

static inline int sgn(uint8_t b) { return b ? 1 : -1; } /* utils.hh:79-81 */

/* problem.cc:104-142: E = sum_i h_i s_i + sum_{i<j} w_ij s_i s_j; every
 * undirected entry is stored twice, so the pair sum is half the CSR sum. */
int64_t or_energy(const or_graph* g, const uint8_t* s) {
    int64_t e2 = 0, lin = 0;
    for (int32_t i = 0; i < g->n; ++i) {
        int si = sgn(s[i]);
        lin += (int64_t)g->h[i] * si;
        for (int32_t p = g->rowptr[i]; p < g->rowptr[i + 1]; ++p)
            e2 += (int64_t)g->w[p] * si * sgn(s[g->col[p]]);
    }
    return lin + e2 / 2;
}

/* problem.cc:145-162: cut = sum over edges of w [s_i != s_j]. */
int64_t or_maxcut(const or_graph* g, const uint8_t* s) {
    int64_t c2 = 0;
    for (int32_t i = 0; i < g->n; ++i)
        for (int32_t p = g->rowptr[i]; p < g->rowptr[i + 1]; ++p)
            if (s[i] != s[g->col[p]]) c2 += g->w[p];
    return c2 / 2;
}

/* problem.cc:33-63: dE_i = -2 h_i s_i - 2 s_i sum_j w_ij s_j. */
void or_delta_e(const or_graph* g, const uint8_t* s, int64_t* dE) {
    for (int32_t i = 0; i < g->n; ++i) {
        int si = sgn(s[i]);
        int64_t f = g->h[i];
        for (int32_t p = g->rowptr[i]; p < g->rowptr[i + 1]; ++p) f += (int64_t)g->w[p] * sgn(s[g->col[p]]);
        dE[i] = -2 * si * f;
    }
}

/* problem.cc:164-181, called before the spin flips (replica.cc:20). */
void or_incremental_de(const or_graph* g, int64_t* dE, const uint8_t* s, int32_t index) {
    dE[index] = -dE[index];
    int si = sgn(s[index]);
    for (int32_t p = g->rowptr[index]; p < g->rowptr[index + 1]; ++p) {
        int32_t j = g->col[p];
        dE[j] += 4 * (int64_t)si * sgn(s[j]) * g->w[p];
    }
}

/* DSatur colouring (Brelaz 1979): repeatedly colour the uncoloured vertex
 * with the most distinct colours among its neighbours (ties: larger degree,
 * then smaller index) with the smallest colour none of its neighbours has.
 * Every adjacency entry counts as an edge (zero weights included, as the
 * cluster moves follow them).  A binary max-heap with lazy deletion holds
 * (saturation, degree, vertex); the key order is total, so the result does
 * not depend on the heap. */
typedef struct {
    int32_t sat, deg, v;
} or_dsat;

static int or_dsat_better(or_dsat a, or_dsat b) {
    if (a.sat != b.sat) return a.sat > b.sat;
    if (a.deg != b.deg) return a.deg > b.deg;
    return a.v < b.v;
}

static void or_heap_push(or_dsat* h, int64_t* len, or_dsat e) {
    int64_t i = (*len)++;
    h[i] = e;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!or_dsat_better(h[i], h[p])) break;
        or_dsat t = h[i];
        h[i] = h[p];
        h[p] = t;
        i = p;
    }
}

static or_dsat or_heap_pop(or_dsat* h, int64_t* len) {
    or_dsat top = h[0];
    h[0] = h[--(*len)];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, b = i;
        if (l < *len && or_dsat_better(h[l], h[b])) b = l;
        if (r < *len && or_dsat_better(h[r], h[b])) b = r;
        if (b == i) break;
        or_dsat t = h[i];
        h[i] = h[b];
        h[b] = t;
        i = b;
    }
    return top;
}

int32_t or_colour(const or_graph* g, int32_t* colour) {
    int32_t n = g->n, ncls = 0;
    int32_t maxdeg = 0;
    for (int32_t i = 0; i < n; ++i)
        if (g->rowptr[i + 1] - g->rowptr[i] > maxdeg) maxdeg = g->rowptr[i + 1] - g->rowptr[i];
    const int32_t words = (maxdeg + 2 + 63) / 64;
    uint64_t* nb = (uint64_t*)calloc((size_t)n * words, sizeof(uint64_t)); /* neighbour colours */
    int32_t* sat = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    const int64_t cap = (int64_t)n + g->rowptr[n] + 1;
    or_dsat* heap = (or_dsat*)malloc((size_t)cap * sizeof(or_dsat));
    int64_t len = 0;
    for (int32_t i = 0; i < n; ++i) {
        colour[i] = -1;
        or_dsat e = {0, g->rowptr[i + 1] - g->rowptr[i], i};
        or_heap_push(heap, &len, e);
    }
    for (int32_t k = 0; k < n; ++k) {
        or_dsat e;
        do e = or_heap_pop(heap, &len);
        while (colour[e.v] >= 0 || e.sat != sat[e.v]);
        const int32_t v = e.v;
        int32_t c = 0;
        while (nb[(size_t)v * words + c / 64] >> (c % 64) & 1u) ++c;
        colour[v] = c;
        if (c + 1 > ncls) ncls = c + 1;
        for (int32_t p = g->rowptr[v]; p < g->rowptr[v + 1]; ++p) {
            const int32_t j = g->col[p];
            uint64_t* w = &nb[(size_t)j * words + c / 64];
            if (colour[j] >= 0 || (*w >> (c % 64) & 1u)) continue;
            *w |= (uint64_t)1 << (c % 64);
            ++sat[j];
            or_dsat f = {sat[j], g->rowptr[j + 1] - g->rowptr[j], j};
            or_heap_push(heap, &len, f);
        }
    }
    free(heap);
    free(sat);
    free(nb);
    /* balancing pass: a class above cap = ceil(n / ncls) gives vertices, in
       index order, to the lowest-numbered class with room and none of their
       neighbours (the product's colouring does the same) */
    {
        const int32_t cap = (n + ncls - 1) / ncls;
        int32_t* size = (int32_t*)calloc((size_t)ncls, sizeof(int32_t));
        uint8_t* nbc = (uint8_t*)calloc((size_t)ncls, 1);
        for (int32_t i = 0; i < n; ++i) size[colour[i]]++;
        for (int32_t c = 0; c < ncls; ++c) {
            for (int32_t v = 0; v < n && size[c] > cap; ++v) {
                if (colour[v] != c) continue;
                for (int32_t p = g->rowptr[v]; p < g->rowptr[v + 1]; ++p) nbc[colour[g->col[p]]] = 1;
                for (int32_t c2 = 0; c2 < ncls; ++c2) {
                    if (c2 != c && size[c2] < cap && !nbc[c2]) {
                        colour[v] = c2;
                        size[c]--;
                        size[c2]++;
                        break;
                    }
                }
                for (int32_t p = g->rowptr[v]; p < g->rowptr[v + 1]; ++p) nbc[colour[g->col[p]]] = 0;
            }
        }
        free(nbc);
        free(size);
    }
    return ncls;
}

void or_positions(const or_graph* g, const int32_t* colour, int32_t ncls, int32_t* pos2orig,
                  int32_t* class_start) {
    for (int32_t c = 0; c <= ncls; ++c) class_start[c] = 0;
    for (int32_t i = 0; i < g->n; ++i) class_start[colour[i] + 1]++;
    for (int32_t c = 0; c < ncls; ++c) class_start[c + 1] += class_start[c];
    int32_t* fill = (int32_t*)calloc((size_t)ncls + 1, sizeof(int32_t));
    for (int32_t i = 0; i < g->n; ++i) pos2orig[class_start[colour[i]] + fill[colour[i]]++] = i;
    free(fill);
}

/* ------------------------------------------------------------------------ */
/* Ladders                                                                   */
/* ------------------------------------------------------------------------ */
/* mpi_utils.cc:6-22 */
void or_beta_linspace(double* temps, double T0, double T1, int64_t num_temp, int inverse) {
    double step, curr;
    if (inverse) {
        step = T1 < T0 ? (T0 - T1) / (double)(num_temp - 1) : (T1 - T0) / (double)(num_temp - 1);
        curr = T1 < T0 ? T1 : T0;
    } else {
        step = T1 < T0 ? (1 / T1 - 1 / T0) / (double)(num_temp - 1) : (1 / T0 - 1 / T1) / (double)(num_temp - 1);
        curr = T1 < T0 ? 1 / T0 : 1 / T1;
    }
    for (int64_t i = 0; i < num_temp; ++i) {
        temps[i] = curr;
        curr += step;
    }
}

/* initialization.cc:158-164 */
int32_t or_icm_count(const double* beta, int32_t nt, double char_scale) {
    int32_t nc = 0;
    for (int32_t i = nt - 1; i >= 0; --i) {
        if (beta[i] < char_scale) break;
        nc++;
    }
    return nc;
}

int32_t or_fmax(const or_graph* g) {
    int64_t fmax = 0;
    for (int32_t i = 0; i < g->n; ++i) {
        int64_t f = llabs((long long)g->h[i]);
        for (int32_t p = g->rowptr[i]; p < g->rowptr[i + 1]; ++p) f += llabs((long long)g->w[p]);
        if (f > fmax) fmax = f;
    }
    return (int32_t)fmax;
}

/* thr[t][m-1] = floor(exp(-beta_t * 2m) * 2^32), saturated to 2^32-1: an
 * uphill move with dE = 2m > 0 is accepted iff the 32-bit draw is below it
 * (replica.cc:31-38 with u = draw * 2^-32). */
void or_threshold_table(const double* beta, int32_t nt, int32_t fmax, uint32_t* thr) {
    for (int32_t t = 0; t < nt; ++t)
        for (int32_t m = 1; m <= fmax; ++m) {
            double x = -(beta[t] * (double)(2 * m));
            double p = or_exp(x) * 4294967296.0;
            double fl = floor(p);
            uint32_t v = fl >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)fl;
            thr[(size_t)t * fmax + (m - 1)] = v;
        }
}

/* ------------------------------------------------------------------------ */
/* Engine                                                                    */
/* ------------------------------------------------------------------------ */
struct or_engine {
    or_graph* g;
    int32_t n, nt, chains, nrep, sweeps, nc, fmax, ncls;
    uint64_t seed;
    uint32_t inst, flags;
    double* beta;   /* [chains*nt] */
    uint32_t* thr;  /* [nt][fmax] */
    int32_t* class_start;
    int32_t* pos2orig;
    int32_t* orig2pos;
    uint8_t* spin;  /* [nrep][n] */
    int64_t* E;
    int64_t* bestE;
    int64_t* best_sweep;
    uint8_t* best_state;
    uint64_t* flips;
    int32_t* slot2rep;
    int32_t* rep2slot;
    uint32_t* swap_att;
    uint32_t* swap_succ;
    uint64_t swaps;
    uint64_t step;
    int64_t* dE;        /* reference-order mode */
    uint64_t* ndraw;    /* reference-order mode draw counters */
    int32_t* elog;
    int64_t elog_len, elog_cap;
    int32_t* icm_seed;
    int32_t* icm_size;
    int32_t* icm_gflip;
};

static uint32_t tagof(const or_engine* e, uint32_t tag) {
    return (e->flags & OR_FLAG_PILOT) ? (tag | TAG_PILOT) : tag;
}

static void recompute_energy(or_engine* e, int32_t r) {
    e->E[r] = or_energy(e->g, e->spin + (size_t)r * e->n);
    if (e->dE) or_delta_e(e->g, e->spin + (size_t)r * e->n, e->dE + (size_t)r * e->n);
}

or_engine* or_engine_create(const or_graph* gin, const double* betas, int32_t nt, int32_t chains,
                            uint64_t seed, uint32_t instance, uint32_t flags, int32_t sweeps,
                            int32_t nc_override) {
    if (nt < 1 || chains < 1 || chains > 2 || sweeps < 0) return NULL;
    or_engine* e = (or_engine*)calloc(1, sizeof(or_engine));
    /* private copy of the graph */
    or_graph* g = (or_graph*)calloc(1, sizeof(or_graph));
    *g = *gin;
    int64_t nnz = gin->rowptr[gin->n];
    g->rowptr = (int32_t*)malloc(((size_t)gin->n + 1) * sizeof(int32_t));
    g->col = (int32_t*)malloc((size_t)(nnz + 1) * sizeof(int32_t));
    g->w = (int32_t*)malloc((size_t)(nnz + 1) * sizeof(int32_t));
    g->h = (int32_t*)malloc(((size_t)gin->n + 1) * sizeof(int32_t));
    or_graph_csr(gin, g->rowptr, g->col, g->w, g->h);
    e->g = g;
    e->n = g->n; e->nt = nt; e->chains = chains; e->nrep = nt * chains; e->sweeps = sweeps;
    e->seed = seed; e->inst = instance; e->flags = flags;
    e->beta = (double*)malloc(sizeof(double) * (size_t)e->nrep);
    for (int32_t c = 0; c < chains; ++c) memcpy(e->beta + (size_t)c * nt, betas, sizeof(double) * (size_t)nt);
    e->fmax = or_fmax(g);
    int32_t fm = e->fmax > 0 ? e->fmax : 1;
    e->thr = (uint32_t*)calloc((size_t)nt * fm, sizeof(uint32_t));
    if (e->fmax > 0) or_threshold_table(betas, nt, e->fmax, e->thr);
    int32_t* colour = (int32_t*)malloc(sizeof(int32_t) * ((size_t)e->n + 1));
    e->ncls = or_colour(g, colour);
    e->class_start = (int32_t*)malloc(sizeof(int32_t) * ((size_t)e->ncls + 1));
    e->pos2orig = (int32_t*)malloc(sizeof(int32_t) * ((size_t)e->n + 1));
    e->orig2pos = (int32_t*)malloc(sizeof(int32_t) * ((size_t)e->n + 1));
    or_positions(g, colour, e->ncls, e->pos2orig, e->class_start);
    for (int32_t p = 0; p < e->n; ++p) e->orig2pos[e->pos2orig[p]] = p;
    /* the colour-order sweep is only valid for a proper colouring */
    for (int32_t i = 0; i < e->n; ++i)
        for (int32_t p = g->rowptr[i]; p < g->rowptr[i + 1]; ++p)
            if (colour[g->col[p]] == colour[i]) { free(colour); or_engine_free(e); return NULL; }
    free(colour);
    size_t ns = (size_t)e->nrep * e->n;
    e->spin = (uint8_t*)malloc(ns + 1);
    e->best_state = (uint8_t*)malloc(ns + 1);
    e->E = (int64_t*)calloc((size_t)e->nrep, sizeof(int64_t));
    e->bestE = (int64_t*)calloc((size_t)e->nrep, sizeof(int64_t));
    e->best_sweep = (int64_t*)calloc((size_t)e->nrep, sizeof(int64_t));
    e->flips = (uint64_t*)calloc((size_t)e->nrep, sizeof(uint64_t));
    e->slot2rep = (int32_t*)malloc(sizeof(int32_t) * (size_t)e->nrep);
    e->rep2slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)e->nrep);
    e->swap_att = (uint32_t*)calloc((size_t)e->nrep, sizeof(uint32_t));
    e->swap_succ = (uint32_t*)calloc((size_t)e->nrep, sizeof(uint32_t));
    e->icm_seed = (int32_t*)calloc((size_t)nt, sizeof(int32_t));
    e->icm_size = (int32_t*)calloc((size_t)nt, sizeof(int32_t));
    e->icm_gflip = (int32_t*)calloc((size_t)nt, sizeof(int32_t));
    if (flags & OR_FLAG_REFERENCE_ORDER) {
        e->dE = (int64_t*)calloc(ns + 1, sizeof(int64_t));
        e->ndraw = (uint64_t*)calloc((size_t)e->nrep, sizeof(uint64_t));
    }
    /* initialization.cc:63-76: Bernoulli(1/2) initial bits per replica */
    for (int32_t r = 0; r < e->nrep; ++r) {
        for (int32_t i = 0; i < e->n; ++i) {
            uint32_t x = draw(seed, (uint32_t)i >> 2, (uint32_t)r, 0, tagof(e, TAG_INIT), instance, (uint32_t)i);
            e->spin[(size_t)r * e->n + i] = (uint8_t)(x >> 31);
        }
        recompute_energy(e, r);
        e->bestE[r] = INT64_MAX; /* replica.cc:54: best_ene = DBL_MAX */
        e->best_sweep[r] = 0;
        memcpy(e->best_state + (size_t)r * e->n, e->spin + (size_t)r * e->n, (size_t)e->n);
    }
    for (int32_t s = 0; s < e->nrep; ++s) { e->slot2rep[s] = s; e->rep2slot[s] = s; }
    if (nc_override >= 0) e->nc = nc_override < nt ? nc_override : nt;
    else e->nc = or_icm_count(betas, nt, g->char_scale);
    if (chains < 2) e->nc = 0;
    return e;
}

void or_engine_free(or_engine* e) {
    if (!e) return;
    or_graph_free(e->g);
    free(e->beta); free(e->thr); free(e->class_start); free(e->pos2orig); free(e->orig2pos);
    free(e->spin); free(e->best_state); free(e->E); free(e->bestE); free(e->best_sweep);
    free(e->flips); free(e->slot2rep); free(e->rep2slot); free(e->swap_att); free(e->swap_succ);
    free(e->dE); free(e->ndraw); free(e->elog); free(e->icm_seed); free(e->icm_size);
    free(e->icm_gflip); free(e);
}

static int32_t local_field(const or_graph* g, const uint8_t* s, int32_t i) {
    int32_t f = g->h[i];
    for (int32_t p = g->rowptr[i]; p < g->rowptr[i + 1]; ++p) f += g->w[p] * sgn(s[g->col[p]]);
    return f;
}

/* Colour-class order: classes in order; inside a class every site decides
 * from the state at the class start (no intra-class edges, so sequential and
 * parallel application agree).  Draw for position p: word p&3 of the block
 * p>>2.  Best energy is checked after every class that accepted a flip. */
static void sweep_colour(or_engine* e, int32_t r, uint32_t g_sweep) {
    const or_graph* g = e->g;
    uint8_t* s = e->spin + (size_t)r * e->n;
    int32_t t = e->rep2slot[r] % e->nt;
    const uint32_t* thr = e->thr + (size_t)t * (e->fmax > 0 ? e->fmax : 1);
    for (int32_t c = 0; c < e->ncls; ++c) {
        int64_t dsum = 0;
        uint64_t nacc = 0;
        for (int32_t p = e->class_start[c]; p < e->class_start[c + 1]; ++p) {
            int32_t i = e->pos2orig[p];
            int32_t dE = -2 * sgn(s[i]) * local_field(g, s, i);
            int accept = dE <= 0;
            if (!accept) {
                uint32_t x = draw(e->seed, (uint32_t)p >> 2, (uint32_t)r, g_sweep, tagof(e, TAG_SWEEP), e->inst,
                                  (uint32_t)p);
                accept = x < thr[dE / 2 - 1];
            }
            if (accept) {
                s[i] ^= 1u;
                dsum += dE;
                nacc++;
            }
        }
        e->E[r] += dsum;
        e->flips[r] += nacc;
        if (nacc && e->E[r] < e->bestE[r]) {
            e->bestE[r] = e->E[r];
            e->best_sweep[r] = g_sweep;
            memcpy(e->best_state + (size_t)r * e->n, s, (size_t)e->n);
        }
    }
}

/* Index order, maintained dE, a draw only for uphill proposals, best checked
 * after every accepted flip: replica.cc:5-38 and problem.cc:164-181. */
static void sweep_reference(or_engine* e, int32_t r, uint32_t g_sweep) {
    const or_graph* g = e->g;
    uint8_t* s = e->spin + (size_t)r * e->n;
    int64_t* dE = e->dE + (size_t)r * e->n;
    int32_t t = e->rep2slot[r] % e->nt;
    const uint32_t* thr = e->thr + (size_t)t * (e->fmax > 0 ? e->fmax : 1);
    for (int32_t i = 0; i < e->n; ++i) {
        int accept = dE[i] <= 0;
        if (!accept) {
            uint64_t k = e->ndraw[r]++;
            uint32_t x = draw(e->seed, (uint32_t)(k >> 2), (uint32_t)r, (uint32_t)(k >> 34), tagof(e, TAG_REFSWEEP),
                              e->inst, (uint32_t)k);
            accept = x < thr[dE[i] / 2 - 1];
        }
        if (accept) {
            e->E[r] += dE[i];
            or_incremental_de(g, dE, s, i);
            s[i] ^= 1u;
            if (e->E[r] < e->bestE[r]) {
                e->bestE[r] = e->E[r];
                e->best_sweep[r] = g_sweep;
                memcpy(e->best_state + (size_t)r * e->n, s, (size_t)e->n);
            }
            e->flips[r]++;
        }
    }
}

/* Random permutation of 0..n-1 by sorting random keys (ties by index). */
static void key_permutation(const or_engine* e, int32_t n, uint32_t c1, uint32_t c2, uint32_t tag, int32_t* out) {
    uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    for (int32_t j = 0; j < n; ++j) key[j] = draw(e->seed, (uint32_t)j >> 2, c1, c2, tag, e->inst, (uint32_t)j);
    for (int32_t j = 0; j < n; ++j) {
        int32_t rank = 0;
        for (int32_t i = 0; i < n; ++i)
            if (key[i] < key[j] || (key[i] == key[j] && i < j)) rank++;
        out[rank] = j;
    }
    free(key);
}

static int exchange_accept(const or_engine* e, int32_t k, uint32_t x) {
    int32_t r1 = e->slot2rep[k], r2 = e->slot2rep[k + 1];
    double dB = e->beta[k] - e->beta[k + 1];
    double dE = (double)(e->E[r1] - e->E[r2]);
    double lb = dB * dE;
    return or_uniform_log(x) < lb;
}

static void do_swap(or_engine* e, int32_t k) {
    int32_t r1 = e->slot2rep[k], r2 = e->slot2rep[k + 1];
    e->slot2rep[k] = r2; e->slot2rep[k + 1] = r1;
    e->rep2slot[r2] = k; e->rep2slot[r1] = k + 1;
}

/* replica_exchange.cc:5-100 driven by solver.cc:81-89.  Attempt a (0 <=
 * a < 2(Nt-1)) is chain a&1 and uses entry a mod (Nt-1) of the pair order of
 * half a/(Nt-1); chain 1 reuses chain 0's order (replica_exchange.cc:11-15).
 * Accept iff log(u) < (B_k - B_k+1)(E_k - E_k+1) (:24-40). */
static void exchange_main(or_engine* e, uint32_t s) {
    int32_t nt = e->nt, np = nt - 1;
    if (np <= 0) return;
    int32_t* perm[2];
    perm[0] = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
    perm[1] = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
    key_permutation(e, np, 0, s, tagof(e, TAG_SHUF), perm[0]);
    key_permutation(e, np, 1, s, tagof(e, TAG_SHUF), perm[1]);
    for (int32_t a = 0; a < 2 * np; ++a) {
        int32_t chain = a & 1;
        if (chain >= e->chains) continue;
        int32_t k = perm[a / np][a % np] + chain * nt;
        uint32_t x = draw(e->seed, (uint32_t)a >> 2, 0, s, tagof(e, TAG_EXCH), e->inst, (uint32_t)a);
        e->swap_att[k]++;
        if (exchange_accept(e, k, x)) {
            e->swap_succ[k]++;
            e->swaps++;
            do_swap(e, k);
        }
    }
    free(perm[0]); free(perm[1]);
}

/* cae_estimator.cc:122-139: shuffle the pair order, then try every pair. */
static void exchange_pilot(or_engine* e, uint32_t s) {
    int32_t np = e->nt - 1;
    if (np <= 0) return;
    int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * (size_t)np);
    key_permutation(e, np, 0, s, tagof(e, TAG_SHUF), perm);
    for (int32_t a = 0; a < np; ++a) {
        int32_t k = perm[a];
        uint32_t x = draw(e->seed, (uint32_t)a >> 2, 0, s, tagof(e, TAG_EXCH), e->inst, (uint32_t)a);
        e->swap_att[k]++;
        if (exchange_accept(e, k, x)) {
            e->swap_succ[k]++;
            e->swaps++;
            do_swap(e, k);
        }
    }
    free(perm);
}

/* icm.cc:8-72 and the single-rank branch of icm.cc:142-150,193-212. */
static void icm_pair(or_engine* e, int32_t t, uint32_t s) {
    int32_t n = e->n;
    int32_t r1 = e->slot2rep[t], r2 = e->slot2rep[t + e->nt];
    uint8_t* s1 = e->spin + (size_t)r1 * n;
    uint8_t* s2 = e->spin + (size_t)r2 * n;
    int ref = (e->flags & OR_FLAG_REFERENCE_ORDER) != 0;
    e->icm_seed[t] = -1; e->icm_size[t] = 0; e->icm_gflip[t] = 0;
    int64_t cnt = 0;
    for (int32_t i = 0; i < n; ++i) cnt += s1[i] != s2[i];
    if (cnt > n / 2) { /* icm.cc:20-28 */
        for (int32_t i = 0; i < n; ++i) s1[i] ^= 1u;
        cnt = n - cnt;
        e->icm_gflip[t] = 1;
    }
    if (cnt > 0) {
        uint32_t x = draw(e->seed, (uint32_t)t >> 2, 0, s, tagof(e, TAG_ICM), e->inst, (uint32_t)t);
        int64_t k = (int64_t)(((uint64_t)x * (uint64_t)cnt) >> 32);
        /* k-th element of D: position order for the colour engine, index order
         * for the reference-order engine */
        int32_t seedv = -1;
        int64_t seen = 0;
        for (int32_t q = 0; q < n && seedv < 0; ++q) {
            int32_t i = ref ? q : e->pos2orig[q];
            if (s1[i] != s2[i]) {
                if (seen == k) seedv = i;
                seen++;
            }
        }
        e->icm_seed[t] = ref ? seedv : e->orig2pos[seedv];
        /* icm.cc:37-72: explicit stack, grow the D-component of the seed */
        uint8_t* added = (uint8_t*)calloc((size_t)n, 1);
        int32_t* stack = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
        int32_t sp = 0, size = 0;
        stack[sp++] = seedv;
        added[seedv] = 1;
        while (sp > 0) {
            int32_t i = stack[--sp];
            s1[i] ^= 1u; s2[i] ^= 1u;
            size++;
            for (int32_t p = e->g->rowptr[i]; p < e->g->rowptr[i + 1]; ++p) {
                int32_t j = e->g->col[p];
                if (s1[j] != s2[j] && !added[j]) { added[j] = 1; stack[sp++] = j; }
            }
        }
        e->icm_size[t] = size;
        free(added); free(stack);
    }
    /* icm.cc:193-211: both states written back, E and dE recomputed */
    recompute_energy(e, r1);
    recompute_energy(e, r2);
}

void or_engine_step(or_engine* e, int64_t nsteps) {
    for (int64_t it = 0; it < nsteps; ++it) {
        uint32_t s = (uint32_t)e->step;
        for (int32_t sw = 0; sw < e->sweeps; ++sw) {
            uint32_t gs = (uint32_t)(e->step * (uint64_t)e->sweeps + (uint64_t)sw);
            for (int32_t r = 0; r < e->nrep; ++r) {
                if (e->flags & OR_FLAG_REFERENCE_ORDER) sweep_reference(e, r, gs);
                else sweep_colour(e, r, gs);
            }
        }
        if (e->flags & OR_FLAG_LOG_ENERGY) {
            if (e->elog_len + e->nrep > e->elog_cap) {
                e->elog_cap = (e->elog_cap + e->nrep) * 2;
                e->elog = (int32_t*)realloc(e->elog, sizeof(int32_t) * (size_t)e->elog_cap);
            }
            for (int32_t sl = 0; sl < e->nrep; ++sl) e->elog[e->elog_len + sl] = (int32_t)e->E[e->slot2rep[sl]];
            e->elog_len += e->nrep;
        }
        if (!(e->flags & OR_FLAG_NO_EXCHANGE)) {
            if (e->flags & OR_FLAG_PILOT) exchange_pilot(e, s);
            else exchange_main(e, s);
        }
        if ((e->flags & OR_FLAG_ICM) && e->chains == 2) {
            for (int32_t t = e->nt - 1; t >= e->nt - e->nc; --t) icm_pair(e, t, s);
        }
        e->step++;
    }
}

int32_t or_engine_nrep(const or_engine* e) { return e->nrep; }
int32_t or_engine_nc(const or_engine* e) { return e->nc; }
void or_engine_spins(const or_engine* e, uint8_t* out) { memcpy(out, e->spin, (size_t)e->nrep * e->n); }
void or_engine_set_spins(or_engine* e, const uint8_t* in) {
    memcpy(e->spin, in, (size_t)e->nrep * e->n);
    for (int32_t r = 0; r < e->nrep; ++r) recompute_energy(e, r);
}
void or_engine_energies(const or_engine* e, int64_t* out) { memcpy(out, e->E, sizeof(int64_t) * (size_t)e->nrep); }
void or_engine_slot2rep(const or_engine* e, int32_t* out) { memcpy(out, e->slot2rep, sizeof(int32_t) * (size_t)e->nrep); }
void or_engine_swap_stats(const or_engine* e, uint32_t* att, uint32_t* succ, uint64_t* swaps) {
    memcpy(att, e->swap_att, sizeof(uint32_t) * (size_t)e->nrep);
    memcpy(succ, e->swap_succ, sizeof(uint32_t) * (size_t)e->nrep);
    *swaps = e->swaps;
}
void or_engine_flips(const or_engine* e, uint64_t* out) { memcpy(out, e->flips, sizeof(uint64_t) * (size_t)e->nrep); }
void or_engine_best(const or_engine* e, int64_t* bestE, int64_t* best_sweep, uint8_t* best_state) {
    memcpy(bestE, e->bestE, sizeof(int64_t) * (size_t)e->nrep);
    memcpy(best_sweep, e->best_sweep, sizeof(int64_t) * (size_t)e->nrep);
    memcpy(best_state, e->best_state, (size_t)e->nrep * e->n);
}
int64_t or_engine_log_len(const or_engine* e) { return e->elog_len; }
void or_engine_energy_log(const or_engine* e, int32_t* out) { memcpy(out, e->elog, sizeof(int32_t) * (size_t)e->elog_len); }
void or_engine_icm_trace(const or_engine* e, int32_t* seed, int32_t* size, int32_t* gflip) {
    memcpy(seed, e->icm_seed, sizeof(int32_t) * (size_t)e->nt);
    memcpy(size, e->icm_size, sizeof(int32_t) * (size_t)e->nt);
    memcpy(gflip, e->icm_gflip, sizeof(int32_t) * (size_t)e->nt);
}

/* ------------------------------------------------------------------------ */
/* Multihistogram DOS and ladder: src/cae_estimator.cc, FP64 with libm.      */
/* ------------------------------------------------------------------------ */
static inline size_t to_bin(double e, double emin, double estep) { /* :11-13 */
    return (size_t)floor((e - emin) / estep);
}
static inline double logplusexp(double x1, double l2) { /* :19-26 */
    double l1 = log(x1);
    if (l1 < l2) { double t = l1; l1 = l2; l2 = t; }
    double x = exp(l2 - l1);
    return l1 + log1p(x);
}
static inline double logplusexp2(double l1, double l2) { /* :28-35 */
    if (l1 < l2) { double t = l1; l1 = l2; l2 = t; }
    double x = exp(l2 - l1);
    return l1 + log1p(x);
}

int64_t or_make_histogram(const int32_t* elog, int64_t n, int64_t stride, double estep, double* emin_out,
                          double* emax_out, uint32_t* hist, int64_t cap) {
    double emin = DBL_MAX, emax = -DBL_MAX;
    for (int64_t i = 0; i < n; ++i) {
        double e = (double)elog[i * stride];
        if (e < emin) emin = e;
        if (e > emax) emax = e;
    }
    int64_t nbins = (int64_t)ceil((emax - emin) / estep) + 1;
    *emin_out = emin; *emax_out = emax;
    if (nbins > cap) return -nbins;
    for (int64_t b = 0; b < nbins; ++b) hist[b] = 0;
    for (int64_t i = 0; i < n; ++i) hist[to_bin((double)elog[i * stride], emin, estep)]++;
    return nbins;
}

double or_partition_func(const double* logdos, int64_t nb, double beta, double emin, double estep) { /* :155-170 */
    double logz = 1;
    int first = 1;
    for (int64_t i = 0; i < nb; ++i) {
        double e = (double)i * estep + emin;
        double logboltz = -beta * e;
        if (first) { first = 0; logz = logboltz + logdos[i]; }
        else logz = logplusexp2(logz, logboltz + logdos[i]);
    }
    return logz;
}

/* :172-226 */
static void calculate_dos(uint32_t* const* hist, const int64_t* hsize, int32_t nh, const double* Z,
                          const double* emin_v, const double* beta_h, double estep, double min_emin,
                          double max_emax, int64_t n, double* logdos, int64_t nbins) {
    double logn = log((double)n);
    double e = min_emin;
    for (int64_t i = 0; i < nbins; ++i, e += estep) {
        double lognum = 1, logdenom = 1;
        int firstnum = 1, firstdenom = 1;
        for (int32_t j = 0; j < nh; ++j) {
            double term = logn - Z[j] - beta_h[j] * e;
            if (firstdenom) { firstdenom = 0; logdenom = term; }
            else logdenom = logplusexp2(term, logdenom);
            double emin_j = emin_v[j];
            if (e < emin_j) continue;
            double index = (double)to_bin(e, emin_j, estep);
            if (index >= (double)hsize[j]) continue;
            uint32_t c = hist[j][(size_t)index];
            if (firstnum && c) { firstnum = 0; lognum = log((double)c); }
            else if (!firstnum && c) lognum = logplusexp((double)c, lognum);
        }
        logdos[i] = firstnum ? -DBL_MAX : lognum - logdenom;
    }
    (void)max_emax;
}

double or_acceptance_rate(const double* logdos, const double* p1, const double* p2, int64_t nb, double beta_1,
                          double beta_2, double estep, double emin) { /* :262-293 */
    (void)logdos;
    double r = 1;
    int first = 1;
    double e1 = emin, e2;
    for (int64_t i = 0; i < nb; ++i, e1 += estep) {
        double p1i = p1[i];
        for (int64_t j = 0; j <= i; ++j) {
            if (first) { first = 0; r = p1i + p2[j]; }
            else r = logplusexp2(r, p1i + p2[j]);
        }
        e2 = e1 + estep;
        for (int64_t j = i + 1; j < nb; ++j, e2 += estep) {
            double p12 = (beta_1 - beta_2) * (e1 - e2);
            if (first) { first = 0; r = p1i + p2[j] + p12; }
            else r = logplusexp2(r, p1i + p2[j] + p12);
        }
    }
    return exp(r);
}

static void get_pe_vec(const double* logdos, int64_t nb, double* logpE, double emin, double beta, double logZ,
                       double estep) { /* :255-260 */
    double e = emin;
    for (int64_t i = 0; i < nb; ++i, e += estep) logpE[i] = logdos[i] - e * beta - logZ;
}

int32_t or_get_beta_from_logs(const int32_t* elog, int64_t nsamples, int32_t nt, const double* hist_betas,
                              double estep, double step_size, double r, double beta_start, double beta_end,
                              double* betas_out, double* logdos_out, int64_t logdos_cap, int64_t* nbins_out,
                              double* logz_out) {
    uint32_t** hist = (uint32_t**)calloc((size_t)nt, sizeof(uint32_t*));
    int64_t* hsize = (int64_t*)calloc((size_t)nt, sizeof(int64_t));
    double* emin_v = (double*)calloc((size_t)nt, sizeof(double));
    double max_emax = -DBL_MAX, min_emin = DBL_MAX;
    for (int32_t j = 0; j < nt; ++j) {
        double emin, emax;
        /* size first */
        int64_t nb = or_make_histogram(elog + j, nsamples, nt, estep, &emin, &emax, NULL, 0);
        nb = -nb;
        hist[j] = (uint32_t*)calloc((size_t)nb + 1, sizeof(uint32_t));
        hsize[j] = or_make_histogram(elog + j, nsamples, nt, estep, &emin, &emax, hist[j], nb);
        emin_v[j] = emin;
        if (emax > max_emax) max_emax = emax;
        if (emin < min_emin) min_emin = emin;
    }
    int64_t n = 0;
    for (int64_t b = 0; b < hsize[0]; ++b) n += hist[0][b];
    int64_t nbins = (int64_t)ceil((max_emax - min_emin) / estep);
    double* logdos = (double*)calloc((size_t)nbins + 1, sizeof(double));
    double* Z = (double*)calloc((size_t)nt, sizeof(double));
    /* get_partition_iterate :228-253 */
    double emin = min_emin;
    for (int32_t j = 1; j < nt; ++j) if (emin_v[j] < emin) emin = emin_v[j];
    double max_diff;
    do {
        max_diff = 0;
        calculate_dos(hist, hsize, nt, Z, emin_v, hist_betas, estep, emin, max_emax, n, logdos, nbins);
        for (int32_t j = 0; j < nt; ++j) {
            double tz = or_partition_func(logdos, nbins, hist_betas[j], emin, estep);
            max_diff += pow(fabs(Z[j] - tz) / fabs(tz), 2);
            Z[j] = tz;
        }
    } while (max_diff > 1e-7);
    if (logz_out) memcpy(logz_out, Z, sizeof(double) * (size_t)nt);
    if (logdos_out) memcpy(logdos_out, logdos, sizeof(double) * (size_t)(nbins < logdos_cap ? nbins : logdos_cap));
    if (nbins_out) *nbins_out = nbins;
    /* ladder search :360-425 */
    betas_out[0] = beta_start;
    double part_last = or_partition_func(logdos, nbins, beta_start, min_emin, estep);
    double step = step_size;
    double* prev = (double*)calloc((size_t)nbins + 1, sizeof(double));
    double* curr = (double*)calloc((size_t)nbins + 1, sizeof(double));
    get_pe_vec(logdos, nbins, prev, min_emin, beta_start, part_last, estep);
    int32_t i;
    for (i = 1; i < nt; ++i) {
        double beta_prev = betas_out[i - 1];
        double beta_curr = beta_prev;
        double part_curr = 0, r_curr = 0, r_last = -1;
        int cont;
        do {
            cont = 0;
            double beta_orig = beta_curr;
            beta_curr = 1 / (1 / beta_curr + step);
            part_curr = or_partition_func(logdos, nbins, beta_curr, min_emin, estep);
            get_pe_vec(logdos, nbins, curr, min_emin, beta_curr, part_curr, estep);
            r_curr = or_acceptance_rate(logdos, prev, curr, nbins, beta_prev, beta_curr, estep, min_emin);
            if ((r - r_curr) > 0.05) {
                beta_curr = beta_orig;
                step /= 2;
                cont = 1;
                continue;
            }
            if (r_last > 0 && fabs(r_curr - r_last) < 5e-2) {
                step *= 2;
                continue;
            }
            r_last = r_curr;
        } while ((r_curr > r && beta_curr > beta_end) || cont);
        betas_out[i] = beta_curr;
        if (beta_curr <= beta_end) break;
        part_last = part_curr;
        double* tmp = prev; prev = curr; curr = tmp;
    }
    for (int32_t j = 0; j < nt; ++j) free(hist[j]);
    free(hist); free(hsize); free(emin_v); free(logdos); free(Z); free(prev); free(curr);
    return i;
}

void or_gather_samples(const or_graph* g, const double* hist_betas, int32_t nt, uint64_t seed, uint32_t instance,
                       int64_t burn_in, int64_t nsamples, int32_t* elog, uint32_t extra_flags) {
    or_engine* e = or_engine_create(g, hist_betas, nt, 1, seed, instance,
                                    OR_FLAG_PILOT | OR_FLAG_LOG_ENERGY | extra_flags, 1, 0);
    if (!e) return;
    or_engine_step(e, burn_in + nsamples);
    memcpy(elog, e->elog + (size_t)burn_in * nt, sizeof(int32_t) * (size_t)(nsamples * nt));
    or_engine_free(e);
}
