// Host CSR / CSC matrices of the continuous-MPS reader.  Same classes,
// constructors and accessors as the reference reader's interface
// (cpp/mps_reader/include/mps_reader/sparse_matrix.hpp:7-61), so callers of
// read_mps compile unchanged.  The constructors validate the compressed
// layout (pointer size, monotone from 0 to nnz, index range) and throw
// std::invalid_argument.
#pragma once

#include <vector>

namespace mps_reader {

class CscMatrix;

class CsrMatrix {
 public:
  CsrMatrix(int rows, int cols, std::vector<int> row_ptr, std::vector<int> col_idx, std::vector<double> values);

  int rows() const noexcept { return rows_; }
  int cols() const noexcept { return cols_; }
  int nnz() const noexcept { return static_cast<int>(values_.size()); }

  const std::vector<int>& row_ptr() const noexcept { return row_ptr_; }
  const std::vector<int>& col_idx() const noexcept { return col_idx_; }
  const std::vector<double>& values() const noexcept { return values_; }

  // Stable: the entries of a column keep their row order.
  CscMatrix to_csc() const;

 private:
  int rows_;
  int cols_;
  std::vector<int> row_ptr_;
  std::vector<int> col_idx_;
  std::vector<double> values_;
};

class CscMatrix {
 public:
  CscMatrix(int rows, int cols, std::vector<int> col_ptr, std::vector<int> row_idx, std::vector<double> values);

  int rows() const noexcept { return rows_; }
  int cols() const noexcept { return cols_; }
  int nnz() const noexcept { return static_cast<int>(values_.size()); }

  const std::vector<int>& col_ptr() const noexcept { return col_ptr_; }
  const std::vector<int>& row_idx() const noexcept { return row_idx_; }
  const std::vector<double>& values() const noexcept { return values_; }

  // Stable: the entries of a row keep their column order.
  CsrMatrix to_csr() const;

 private:
  int rows_;
  int cols_;
  std::vector<int> col_ptr_;
  std::vector<int> row_idx_;
  std::vector<double> values_;
};

}  // namespace mps_reader
