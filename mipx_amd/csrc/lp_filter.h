// lp_filter.h -- host-side drop of near-zero matrix entries.
//
// Reference src/lp/pdlp.cpp:499-518: when matrix_zero_tol > 0 the entries
// with |a| < matrix_zero_tol leave the matrix that is scaled and iterated on
// (row order and the order within a row are kept).  Objective, bounds and
// the infeasibility / unboundedness certificates still use the original LP,
// so callers keep both.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/mipx_pdlp.h"

namespace mpx {

struct FilteredLp {
    std::vector<int32_t> rs, ci;
    std::vector<double> val;
    mpx_lp lp{};

    FilteredLp(const mpx_lp& src, double tol) : lp(src) {
        const int m = src.num_rows;
        rs.resize(static_cast<size_t>(m) + 1);
        ci.reserve(static_cast<size_t>(src.nnz));
        val.reserve(static_cast<size_t>(src.nnz));
        for (int i = 0; i < m; ++i) {
            rs[static_cast<size_t>(i)] = static_cast<int32_t>(val.size());
            for (int32_t k = src.row_starts[i]; k < src.row_starts[i + 1]; ++k)
                if (std::abs(src.values[k]) >= tol) {
                    ci.push_back(src.col_indices[k]);
                    val.push_back(src.values[k]);
                }
        }
        rs[static_cast<size_t>(m)] = static_cast<int32_t>(val.size());
        lp.nnz = static_cast<int32_t>(val.size());
        lp.row_starts = rs.data();
        lp.col_indices = ci.data();
        lp.values = val.data();
    }
    FilteredLp(const FilteredLp&) = delete;
    FilteredLp& operator=(const FilteredLp&) = delete;
};

}  // namespace mpx
