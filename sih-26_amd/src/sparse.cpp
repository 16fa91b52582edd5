// SPDX-License-Identifier: Apache-2.0
// Column-compressed sparse matrix (reference behaviour: src/la/sparse.cpp:22-302).
#include "sankhya/sparse.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace sankhya {

void SparseMatrix::reset(Index num_rows, Index num_cols) {
  num_rows_ = num_rows;
  num_cols_ = num_cols;
  frozen_ = false;
  overflowed_ = false;
  column_starts_.clear();
  row_indices_.clear();
  values_.clear();
  build_rows_.clear();
  build_cols_.clear();
  build_values_.clear();
}

void SparseMatrix::reserve(std::size_t n) {
  n = std::min(n, static_cast<std::size_t>(kMaxNonzeros));
  build_rows_.reserve(n);
  build_cols_.reserve(n);
  build_values_.reserve(n);
}

void SparseMatrix::add_entry(Index row, Index col, double value) {
  assert(!frozen_);
  if (build_values_.size() >= static_cast<std::size_t>(kMaxNonzeros)) {
    overflowed_ = true;
    return;
  }
  build_rows_.push_back(row);
  build_cols_.push_back(col);
  build_values_.push_back(value);
}

void SparseMatrix::finalize(double drop_tol) {
  if (frozen_) return;
  column_starts_.assign(static_cast<std::size_t>(num_cols_) + 1, 0);
  if (overflowed_) {
    build_rows_.clear();
    build_cols_.clear();
    build_values_.clear();
    frozen_ = true;
    return;
  }
  const std::size_t nnz = build_values_.size();
  // Bucket by column (counting sort), then order each column by row and fold duplicates.
  std::vector<Index> counts(static_cast<std::size_t>(num_cols_) + 1, 0);
  for (std::size_t k = 0; k < nnz; ++k) ++counts[static_cast<std::size_t>(build_cols_[k]) + 1];
  for (Index j = 0; j < num_cols_; ++j) counts[j + 1] += counts[j];
  std::vector<std::pair<Index, double>> bucket(nnz);
  {
    std::vector<Index> next(counts.begin(), counts.end() - 1);
    for (std::size_t k = 0; k < nnz; ++k) {
      bucket[static_cast<std::size_t>(next[build_cols_[k]]++)] = {build_rows_[k], build_values_[k]};
    }
  }
  row_indices_.clear();
  values_.clear();
  row_indices_.reserve(nnz);
  values_.reserve(nnz);
  for (Index j = 0; j < num_cols_; ++j) {
    column_starts_[j] = static_cast<Index>(row_indices_.size());
    auto first = bucket.begin() + counts[j];
    auto last = bucket.begin() + counts[j + 1];
    std::stable_sort(first, last, [](const auto& a, const auto& b) { return a.first < b.first; });
    for (auto p = first; p != last;) {
      const Index row = p->first;
      double sum = 0.0;
      while (p != last && p->first == row) {
        sum += p->second;
        ++p;
      }
      // Keep NaN and infinities so validate() can name them.
      if (!(std::fabs(sum) < drop_tol)) {
        row_indices_.push_back(row);
        values_.push_back(sum);
      }
    }
  }
  column_starts_[num_cols_] = static_cast<Index>(row_indices_.size());
  build_rows_ = {};
  build_cols_ = {};
  build_values_ = {};
  frozen_ = true;
}

void SparseMatrix::assign_csc(Index num_rows, Index num_cols, std::vector<Index> starts,
                              std::vector<Index> rows, std::vector<double> values) {
  reset(num_rows, num_cols);
  column_starts_ = std::move(starts);
  row_indices_ = std::move(rows);
  values_ = std::move(values);
  frozen_ = true;
}

ColumnView SparseMatrix::column(Index j) const noexcept {
  ColumnView v;
  const Index b = column_starts_[j], e = column_starts_[j + 1];
  v.size = e - b;
  if (v.size > 0) {
    v.rows = row_indices_.data() + b;
    v.values = values_.data() + b;
  }
  return v;
}

void SparseMatrix::multiply(const double* x, double* y) const {
  std::fill(y, y + num_rows_, 0.0);
  for (Index j = 0; j < num_cols_; ++j) {
    const double xj = x[j];
    if (xj == 0.0) continue;
    for (Index k = column_starts_[j]; k < column_starts_[j + 1]; ++k) y[row_indices_[k]] += values_[k] * xj;
  }
}

void SparseMatrix::transpose_multiply(const double* x, double* y) const {
  for (Index j = 0; j < num_cols_; ++j) {
    double dot = 0.0;
    for (Index k = column_starts_[j]; k < column_starts_[j + 1]; ++k) dot += values_[k] * x[row_indices_[k]];
    y[j] = dot;
  }
}

}  // namespace sankhya
