// sparse_matrix.cpp -- CSR matrix of the host mirror (see sparse_matrix.h).
#include "mipx/sparse_matrix.h"

#include <algorithm>
#include <cassert>

namespace mipx {

SparseMatrix::SparseMatrix(Index rows, Index cols)
    : rows_(rows), cols_(cols), row_starts_(static_cast<size_t>(rows) + 1, 0) {}

SparseMatrix::SparseMatrix(Index rows, Index cols, std::vector<Triplet> t)
    : rows_(rows), cols_(cols) {
    std::stable_sort(t.begin(), t.end(), [](const Triplet& a, const Triplet& b) {
        return a.row != b.row ? a.row < b.row : a.col < b.col;
    });
    row_starts_.assign(static_cast<size_t>(rows) + 1, 0);
    size_t k = 0;
    while (k < t.size()) {
        const Index r = t[k].row, c = t[k].col;
        Real sum = 0.0;
        while (k < t.size() && t[k].row == r && t[k].col == c) sum += t[k++].val;
        if (sum != 0.0) {
            values_.push_back(sum);
            col_indices_.push_back(c);
            ++row_starts_[static_cast<size_t>(r) + 1];
        }
    }
    for (Index i = 0; i < rows; ++i) row_starts_[i + 1] += row_starts_[i];
}

SparseMatrix::SparseMatrix(Index rows, Index cols, std::vector<Real> values,
                           std::vector<Index> col_indices, std::vector<Index> row_starts)
    : rows_(rows),
      cols_(cols),
      values_(std::move(values)),
      col_indices_(std::move(col_indices)),
      row_starts_(std::move(row_starts)) {}

void SparseMatrix::multiply(std::span<const Real> x, std::span<Real> y) const {
    for (Index i = 0; i < rows_; ++i) {
        Real s = 0.0;
        for (Index k = row_starts_[i]; k < row_starts_[i + 1]; ++k) s += values_[k] * x[col_indices_[k]];
        y[i] = s;
    }
}

void SparseMatrix::multiplyTranspose(std::span<const Real> x, std::span<Real> y) const {
    std::fill(y.begin(), y.end(), 0.0);
    for (Index i = 0; i < rows_; ++i) {
        if (x[i] == 0.0) continue;
        for (Index k = row_starts_[i]; k < row_starts_[i + 1]; ++k) y[col_indices_[k]] += values_[k] * x[i];
    }
}

void SparseMatrix::addRow(std::span<const Index> indices, std::span<const Real> values) {
    std::vector<std::pair<Index, Real>> e;
    for (size_t k = 0; k < indices.size(); ++k) e.emplace_back(indices[k], values[k]);
    std::sort(e.begin(), e.end());
    for (auto& [c, v] : e) {
        col_indices_.push_back(c);
        values_.push_back(v);
    }
    row_starts_.push_back(static_cast<Index>(values_.size()));
    ++rows_;
}

void SparseMatrix::removeRowStable(Index i) {
    const Index b = row_starts_[i], e = row_starts_[i + 1], len = e - b;
    values_.erase(values_.begin() + b, values_.begin() + e);
    col_indices_.erase(col_indices_.begin() + b, col_indices_.begin() + e);
    row_starts_.erase(row_starts_.begin() + i + 1);
    for (size_t r = static_cast<size_t>(i) + 1; r < row_starts_.size(); ++r) row_starts_[r] -= len;
    --rows_;
}

}  // namespace mipx
