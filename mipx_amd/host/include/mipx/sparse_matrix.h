// sparse_matrix.h -- CSR matrix as PDLP consumes it.
// Interface follows the reference's SparseMatrix (include/mipx/sparse_matrix.h:27-106):
// triplet and CSR constructors, multiply / multiplyTranspose and raw CSR
// accessors, plus the row edits PdlpSolver's mutators need.
#pragma once

#include <span>
#include <vector>

#include "mipx/core.h"

namespace mipx {

struct Triplet {
    Index row;
    Index col;
    Real val;
};

class SparseMatrix {
public:
    SparseMatrix(Index rows, Index cols);
    // Sorts by (row, col), sums duplicates, drops entries that sum to zero.
    SparseMatrix(Index rows, Index cols, std::vector<Triplet> triplets);
    // Pre-formed CSR (column indices sorted within each row).
    SparseMatrix(Index rows, Index cols, std::vector<Real> values, std::vector<Index> col_indices,
                 std::vector<Index> row_starts);

    [[nodiscard]] Index numRows() const { return rows_; }
    [[nodiscard]] Index numCols() const { return cols_; }
    [[nodiscard]] Index numNonzeros() const { return static_cast<Index>(values_.size()); }

    void multiply(std::span<const Real> x, std::span<Real> y) const;           // y = A x
    void multiplyTranspose(std::span<const Real> x, std::span<Real> y) const;  // y = A' x

    void addRow(std::span<const Index> indices, std::span<const Real> values);
    void removeRowStable(Index i);

    [[nodiscard]] std::span<const Real> csr_values() const { return values_; }
    [[nodiscard]] std::span<const Index> csr_col_indices() const { return col_indices_; }
    [[nodiscard]] std::span<const Index> csr_row_starts() const { return row_starts_; }

private:
    Index rows_;
    Index cols_;
    std::vector<Real> values_;
    std::vector<Index> col_indices_;
    std::vector<Index> row_starts_;
};

}  // namespace mipx
