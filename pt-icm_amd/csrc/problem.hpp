// problem.hpp -- host-side instance model: gset/.ising reader, integer CSR,
// energy/maxcut/dE, DSatur colouring and the colour-sorted position order the
// kernels sweep in.  Restates pticm::Problem (include/problem.hh:7-98,
// src/problem.cc:1-247) with integer weights.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ptx {

struct Status {
    int code = 0;  // 0 ok, otherwise a PTX_ERR_* value
    std::string msg;
    static Status ok() { return {}; }
    static Status err(int c, std::string m) { return {c, std::move(m)}; }
    explicit operator bool() const { return code == 0; }
};

struct Problem {
    int32_t n = 0;
    int64_t m_header = 0;
    bool sparse = true;
    double char_scale = 0.0;
    // adjacency in index order (each undirected entry stored at both ends)
    std::vector<int32_t> rowptr, col, w, h;
    // colour classes (DSatur) and the position order
    int32_t ncls = 0;
    std::vector<int32_t> colour, pos2orig, orig2pos, class_start;
    int32_t fmax = 0;
    int32_t max_deg = 0;
    int64_t abs_sum = 0;  // sum over edges |w| + sum |h|
    bool has_field = false;
    // periodic hypercubic lattice in index order (i = x + L*y [+ L*L*z]),
    // every row exactly the 2*lat_dim wrapped axis neighbours: cluster moves
    // may then grow clusters as bit sets (lat_dim 0: not such a lattice)
    int32_t lat_dim = 0, lat_L = 0;

    int64_t nnz() const { return rowptr.empty() ? 0 : rowptr[n]; }
    int64_t energy(const uint8_t* s) const;
    int64_t maxcut(const uint8_t* s) const;
    int64_t delta_e(const uint8_t* s, int64_t* dE) const;
    Status finalize();  // colouring, positions, bounds, limits, lattice detection
    void detect_lattice();
};

// src/problem.cc:183-247
Status read_gset(const std::string& path, Problem& out);
// include/problem.hh:40-62 (sparse) and src/problem.cc:210-215 (dense)
Status build_from_coo(int32_t n, int64_t m, const int32_t* src, const int32_t* dst, const double* w,
                      const double* h, bool dense, Problem& out);

}  // namespace ptx
