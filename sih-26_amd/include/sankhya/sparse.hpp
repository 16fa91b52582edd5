// SPDX-License-Identifier: Apache-2.0
// Column-compressed sparse matrix: build from triplets, then freeze (reference:
// include/sankhya/sparse.hpp, src/la/sparse.cpp:22-309). After finalize() the row indices of
// every column ascend and duplicates are summed, which is the order the PDHG path's
// scaling and transposes rely on.
#pragma once

#include <cassert>
#include <cstddef>
#include <vector>

#include "sankhya/tolerances.hpp"
#include "sankhya/types.hpp"

namespace sankhya {

struct ColumnView {
  Index size = 0;
  const Index* rows = nullptr;
  const double* values = nullptr;
};

class SparseMatrix {
 public:
  SparseMatrix() = default;
  SparseMatrix(Index num_rows, Index num_cols) { reset(num_rows, num_cols); }

  void reset(Index num_rows, Index num_cols);
  void reserve(std::size_t n);
  void add_entry(Index row, Index col, double value);
  /// Sort each column by row, sum duplicates, drop |sum| < drop_tol (NaN kept).
  void finalize(double drop_tol = tol::kZeroDrop);

  [[nodiscard]] bool frozen() const noexcept { return frozen_; }
  [[nodiscard]] bool overflowed() const noexcept { return overflowed_; }
  [[nodiscard]] Index nonzero_limit() const noexcept { return kMaxNonzeros; }
  [[nodiscard]] Index num_rows() const noexcept { return num_rows_; }
  [[nodiscard]] Index num_cols() const noexcept { return num_cols_; }
  [[nodiscard]] Index num_nonzeros() const noexcept {
    return static_cast<Index>(frozen_ ? values_.size() : build_values_.size());
  }
  [[nodiscard]] const std::vector<Index>& column_starts() const noexcept { return column_starts_; }
  [[nodiscard]] const std::vector<Index>& row_indices() const noexcept { return row_indices_; }
  [[nodiscard]] const std::vector<double>& values() const noexcept { return values_; }
  [[nodiscard]] std::vector<double>& mutable_values() noexcept { return values_; }
  [[nodiscard]] ColumnView column(Index j) const noexcept;

  /// Adopt an already column-compressed pattern (rows ascending per column, no duplicates).
  void assign_csc(Index num_rows, Index num_cols, std::vector<Index> starts,
                  std::vector<Index> rows, std::vector<double> values);

  /// y = A x and y = A^T x on the host (used by the quality report, not the hot path).
  void multiply(const double* x, double* y) const;
  void transpose_multiply(const double* x, double* y) const;

 private:
  Index num_rows_ = 0, num_cols_ = 0;
  bool frozen_ = false, overflowed_ = false;
  std::vector<Index> column_starts_, row_indices_;
  std::vector<double> values_;
  std::vector<Index> build_rows_, build_cols_;
  std::vector<double> build_values_;
};

}  // namespace sankhya
