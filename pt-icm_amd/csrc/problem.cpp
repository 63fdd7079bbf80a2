// problem.cpp -- see problem.hpp.
#include "problem.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <queue>
#include <sstream>

#include "../../include/ptx.h"

namespace ptx {

namespace {

struct RawEdge {
    int64_t u, v;
    double w;
};

bool integral(double x) { return std::isfinite(x) && x == std::floor(x) && std::fabs(x) <= 32767.0; }

// Builds the CSR.  Sparse instances keep every listed entry, duplicates and
// zero weights included, at both ends in file order (problem.cc:226-229; the
// sparse branch of bfs_icm, icm.cc:53-58, follows every listed entry).  Dense
// instances follow the matrix: the last weight of a pair wins and zero
// entries are no edge (problem.cc:230-232, icm.cc:61-66).
Status assemble(int32_t n, int64_t m_header, bool sparse, const std::vector<RawEdge>& edges,
                const std::vector<double>& hin, double char_scale, Problem& P) {
    P = Problem();
    P.n = n;
    P.m_header = m_header;
    P.sparse = sparse;
    P.char_scale = char_scale;
    P.h.assign(n, 0);
    for (int32_t i = 0; i < n; ++i) {
        if (!integral(hin[i]))
            return Status::err(PTX_ERR_UNSUPPORTED, "linear term is not an integer in [-32767, 32767]");
        P.h[i] = (int32_t)hin[i];
    }
    std::vector<int64_t> deg(n, 0);
    for (const auto& e : edges) {
        if (e.u < 0 || e.u >= n || e.v < 0 || e.v >= n) return Status::err(PTX_ERR_INVALID, "edge endpoint out of range");
        if (!integral(e.w)) return Status::err(PTX_ERR_UNSUPPORTED, "edge weight is not an integer in [-32767, 32767]");
    }
    if (sparse) {
        for (const auto& e : edges) {
            if (e.u == e.v) return Status::err(PTX_ERR_UNSUPPORTED, "self-loop in a sparse instance");
            deg[e.u]++;
            deg[e.v]++;
        }
        P.rowptr.assign(n + 1, 0);
        for (int32_t i = 0; i < n; ++i) P.rowptr[i + 1] = P.rowptr[i] + (int32_t)deg[i];
        P.col.assign(P.rowptr[n], 0);
        P.w.assign(P.rowptr[n], 0);
        std::vector<int32_t> fill(P.rowptr.begin(), P.rowptr.end() - 1);
        for (const auto& e : edges) {
            P.col[fill[e.u]] = (int32_t)e.v;
            P.w[fill[e.u]++] = (int32_t)e.w;
            P.col[fill[e.v]] = (int32_t)e.u;
            P.w[fill[e.v]++] = (int32_t)e.w;
        }
    } else {
        // dense: row-major matrix, last write wins
        std::vector<double> mat((size_t)n * n, 0.0);
        for (const auto& e : edges) {
            mat[(size_t)e.u * n + e.v] = e.w;
            mat[(size_t)e.v * n + e.u] = e.w;
        }
        P.rowptr.assign(n + 1, 0);
        for (int32_t i = 0; i < n; ++i) {
            int32_t c = 0;
            for (int32_t j = 0; j < n; ++j)
                if (j != i && std::fabs(mat[(size_t)i * n + j]) > 1e-10) ++c;
            P.rowptr[i + 1] = P.rowptr[i] + c;
        }
        P.col.reserve(P.rowptr[n]);
        P.w.reserve(P.rowptr[n]);
        for (int32_t i = 0; i < n; ++i)
            for (int32_t j = 0; j < n; ++j)
                if (j != i && std::fabs(mat[(size_t)i * n + j]) > 1e-10) {
                    P.col.push_back(j);
                    P.w.push_back((int32_t)mat[(size_t)i * n + j]);
                }
    }
    return P.finalize();
}

}  // namespace

Status Problem::finalize() {
    if (n <= 0) return Status::err(PTX_ERR_INVALID, "empty problem");
    if (n > 65536) return Status::err(PTX_ERR_UNSUPPORTED, "more than 65536 spins");
    // DSatur colouring (Brelaz 1979): the uncoloured vertex with the most
    // distinct neighbour colours (ties: larger degree, then smaller index) takes
    // the smallest colour its neighbours lack.  Fewer classes mean fewer
    // dependent phases per sweep (G022: 8 instead of 12 for greedy index
    // order).  A max-heap with lazy deletion; the key order is total.
    max_deg = 0;
    for (int32_t i = 0; i < n; ++i) max_deg = std::max(max_deg, rowptr[i + 1] - rowptr[i]);
    colour.assign(n, -1);
    {
        struct Key {
            int32_t sat, deg, v;
            bool operator<(const Key& o) const {  // "worse than": the heap's top is the best key
                if (sat != o.sat) return sat < o.sat;
                if (deg != o.deg) return deg < o.deg;
                return v > o.v;
            }
        };
        const int32_t words = (max_deg + 2 + 63) / 64;
        std::vector<uint64_t> nb((size_t)n * words, 0);  // neighbour colours per vertex
        std::vector<int32_t> sat(n, 0);
        std::priority_queue<Key> heap;
        for (int32_t i = 0; i < n; ++i) heap.push({0, rowptr[i + 1] - rowptr[i], i});
        ncls = 0;
        for (int32_t k = 0; k < n; ++k) {
            Key e;
            do {
                e = heap.top();
                heap.pop();
            } while (colour[e.v] >= 0 || e.sat != sat[e.v]);
            const int32_t v = e.v;
            int32_t c = 0;
            while ((nb[(size_t)v * words + c / 64] >> (c % 64)) & 1u) ++c;
            colour[v] = c;
            ncls = std::max(ncls, c + 1);
            for (int32_t q = rowptr[v]; q < rowptr[v + 1]; ++q) {
                const int32_t j = col[q];
                uint64_t& w = nb[(size_t)j * words + c / 64];
                if (colour[j] >= 0 || ((w >> (c % 64)) & 1u)) continue;
                w |= (uint64_t)1 << (c % 64);
                heap.push({++sat[j], rowptr[j + 1] - rowptr[j], j});
            }
        }
    }
    // Balancing pass: a class above cap = ceil(n / ncls) gives vertices, in
    // index order, to the lowest-numbered class with room and none of
    // their neighbours (the colouring stays proper, no class empties).
    // Classes of at most half a workgroup's lanes let two lanes share each
    // position in the cooperative sweep.
    {
        const int32_t cap = (n + ncls - 1) / ncls;
        std::vector<int32_t> size(ncls, 0);
        for (int32_t i = 0; i < n; ++i) size[colour[i]]++;
        std::vector<uint8_t> nbc(ncls, 0);
        for (int32_t c = 0; c < ncls; ++c) {
            for (int32_t v = 0; v < n && size[c] > cap; ++v) {
                if (colour[v] != c) continue;
                for (int32_t q = rowptr[v]; q < rowptr[v + 1]; ++q) nbc[colour[col[q]]] = 1;
                for (int32_t c2 = 0; c2 < ncls; ++c2) {
                    if (c2 != c && size[c2] < cap && !nbc[c2]) {
                        colour[v] = c2;
                        size[c]--;
                        size[c2]++;
                        break;
                    }
                }
                for (int32_t q = rowptr[v]; q < rowptr[v + 1]; ++q) nbc[colour[col[q]]] = 0;
            }
        }
    }
    class_start.assign(ncls + 1, 0);
    for (int32_t i = 0; i < n; ++i) class_start[colour[i] + 1]++;
    for (int32_t c = 0; c < ncls; ++c) class_start[c + 1] += class_start[c];
    pos2orig.assign(n, 0);
    orig2pos.assign(n, 0);
    std::vector<int32_t> fill(class_start.begin(), class_start.end() - 1);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t p = fill[colour[i]]++;
        pos2orig[p] = i;
        orig2pos[i] = p;
    }
    // field bound and energy range
    fmax = 0;
    abs_sum = 0;
    has_field = false;
    int64_t wsum2 = 0;
    for (int32_t i = 0; i < n; ++i) {
        int64_t f = std::llabs((long long)h[i]);
        if (h[i] != 0) has_field = true;
        for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) f += std::llabs((long long)w[q]);
        for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) wsum2 += std::llabs((long long)w[q]);
        fmax = (int32_t)std::max<int64_t>(fmax, f);
        abs_sum += std::llabs((long long)h[i]);
    }
    abs_sum += wsum2 / 2;
    if (abs_sum > (1ll << 30)) return Status::err(PTX_ERR_UNSUPPORTED, "energy range exceeds 32-bit engine energies");
    if (fmax > 8192) return Status::err(PTX_ERR_UNSUPPORTED, "local field bound above 8192 (threshold table too large)");
    detect_lattice();
    return Status::ok();
}

void Problem::detect_lattice() {
    lat_dim = 0;
    lat_L = 0;
    for (int32_t D = 2; D <= 3; ++D) {
        int32_t L = (int32_t)std::lround(std::pow((double)n, 1.0 / D));
        int64_t nn = 1;
        for (int32_t d = 0; d < D; ++d) nn *= L;
        if (L < 3 || nn != n) continue;
        bool ok = true;
        for (int32_t i = 0; i < n && ok; ++i) {
            if (rowptr[i + 1] - rowptr[i] != 2 * D) {
                ok = false;
                break;
            }
            // the wrapped neighbours of i along each axis (stride 1, L, L*L)
            int32_t want[6];
            int64_t stride = 1;
            for (int32_t d = 0; d < D; ++d, stride *= L) {
                const int32_t c = (int32_t)((i / stride) % L);
                const int32_t base = i - (int32_t)(c * stride);
                want[2 * d] = base + (int32_t)(((c + 1) % L) * stride);
                want[2 * d + 1] = base + (int32_t)(((c + L - 1) % L) * stride);
            }
            for (int32_t q = rowptr[i]; q < rowptr[i + 1] && ok; ++q) {
                int32_t hits = 0;
                for (int32_t k = 0; k < 2 * D; ++k) hits += want[k] == col[q];
                ok = hits == 1;
            }
            // 2D entries, each naming exactly one wanted neighbour, and the
            // wanted neighbours distinct (L >= 3): the row is that set
            for (int32_t q = rowptr[i]; q < rowptr[i + 1] && ok; ++q)
                for (int32_t q2 = q + 1; q2 < rowptr[i + 1]; ++q2)
                    if (col[q] == col[q2]) ok = false;
        }
        if (ok) {
            lat_dim = D;
            lat_L = L;
            return;
        }
    }
}

int64_t Problem::energy(const uint8_t* s) const {
    int64_t pair2 = 0, lin = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int64_t si = s[i] ? 1 : -1;
        lin += si * h[i];
        int64_t acc = 0;
        for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) acc += (s[col[q]] ? w[q] : -w[q]);
        pair2 += si * acc;
    }
    return lin + pair2 / 2;
}

int64_t Problem::maxcut(const uint8_t* s) const {
    int64_t c2 = 0;
    for (int32_t i = 0; i < n; ++i)
        for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q)
            if (s[i] != s[col[q]]) c2 += w[q];
    return c2 / 2;
}

int64_t Problem::delta_e(const uint8_t* s, int64_t* dE) const {
    for (int32_t i = 0; i < n; ++i) {
        int64_t f = h[i];
        for (int32_t q = rowptr[i]; q < rowptr[i + 1]; ++q) f += (s[col[q]] ? w[q] : -w[q]);
        dE[i] = (s[i] ? -2 : 2) * f;
    }
    return energy(s);
}

Status read_gset(const std::string& path, Problem& out) {
    std::ifstream in(path);
    if (!in.is_open()) return Status::err(PTX_ERR_IO, "cannot open problem file: " + path);
    const bool ising = path.find(".ising") != std::string::npos;
    const bool zero_based = path.find(".gset") != std::string::npos || ising;
    long long N = 0, L = 0, M = 0;
    if (ising) in >> N >> L >> M;
    else in >> N >> M;
    if (!in || N <= 0 || M < 0 || L < 0 || N > 65536) {
        if (N > 65536) return Status::err(PTX_ERR_UNSUPPORTED, "more than 65536 spins");
        return Status::err(PTX_ERR_INVALID, "malformed header in " + path);
    }
    const bool sparse = ((double)M / (double)(N * (N - 1) / 2)) < 0.1;  // problem.cc:198
    const long long off = zero_based ? 0 : -1;                          // problem.cc:208
    std::vector<double> h(N, 0.0);
    std::vector<RawEdge> edges;
    edges.reserve((size_t)M);
    double sum = 0.0, sqsum = 0.0;
    long long nlin = 0;
    std::string line;
    std::getline(in, line);  // remainder of the header line
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() <= 1) continue;  // problem.cc:211
        std::istringstream ls(line);
        if (nlin < L) {
            long long u;
            double w;
            if (!(ls >> u >> w)) return Status::err(PTX_ERR_INVALID, "malformed linear-term line: " + line);
            u += off;
            if (u < 0 || u >= N) return Status::err(PTX_ERR_INVALID, "linear-term index out of range: " + line);
            h[u] = w;
            ++nlin;
            continue;
        }
        long long u, v;
        if (!(ls >> u >> v)) return Status::err(PTX_ERR_INVALID, "malformed edge line: " + line);
        double w = 1.0;  // problem.cc:229-233
        double tmp;
        if (ls >> tmp) w = tmp;
        u += off;
        v += off;
        if (u < 0 || u >= N || v < 0 || v >= N) return Status::err(PTX_ERR_INVALID, "edge endpoint out of range: " + line);
        edges.push_back({u, v, w});
        sum += w;
        sqsum += w * w;
    }
    const double mean = sum / M;
    const double variance = sqsum / M - (sum / M * sum / M);  // problem.cc:244
    const double cs = mean < variance ? variance : mean;     // std::max
    return assemble((int32_t)N, M, sparse, edges, h, cs, out);
}

Status build_from_coo(int32_t n, int64_t m, const int32_t* src, const int32_t* dst, const double* w,
                      const double* h, bool dense, Problem& out) {
    if (n <= 0 || m < 0 || (m > 0 && (!src || !dst || !w))) return Status::err(PTX_ERR_INVALID, "bad COO arguments");
    std::vector<RawEdge> edges((size_t)m);
    double sum = 0.0, sqsum = 0.0;
    for (int64_t e = 0; e < m; ++e) {
        edges[e] = {src[e], dst[e], w[e]};
        sum += w[e];
        sqsum += w[e] * w[e];
    }
    std::vector<double> hv(n, 0.0);
    if (h) std::copy(h, h + n, hv.begin());
    double cs = 0.0;
    if (m > 0) {
        const double mean = sum / m, var = sqsum / m - (sum / m * sum / m);
        cs = mean < var ? var : mean;
    }
    return assemble(n, m, !dense, edges, hv, cs, out);
}

}  // namespace ptx
