// lp_problem.h -- the LP in ranged form, fields as in the reference's
// include/mipx/lp_problem.h:14-64 (the subset the PDLP path reads).
#pragma once

#include <limits>
#include <string>
#include <vector>

#include "mipx/core.h"
#include "mipx/sparse_matrix.h"

namespace mipx {

inline constexpr Real kInf = std::numeric_limits<Real>::infinity();

struct LpProblem {
    std::string name;
    Sense sense = Sense::Minimize;
    Index num_cols = 0;
    std::vector<Real> obj;
    std::vector<Real> col_lower;
    std::vector<Real> col_upper;
    std::vector<VarType> col_type;
    std::vector<std::string> col_names;
    Index num_rows = 0;
    SparseMatrix matrix{0, 0};
    std::vector<Real> row_lower;
    std::vector<Real> row_upper;
    std::vector<std::string> row_names;
    Real obj_offset = 0.0;
};

}  // namespace mipx
