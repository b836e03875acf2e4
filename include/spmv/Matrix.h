// spmv::Matrix<T> -- distributed CSR matrix block (the interface of
// spmv/Matrix.h:23-81) stored and multiplied on the device, backed by the
// spmvx C ABI.  T = double, float, std::complex<double> or
// std::complex<float> (the reference's instantiations, Matrix.cpp:441-444).
#pragma once

#include <memory>

#include "L2GMap.h"
#include "Vector.h"

namespace spmv
{

template <typename T>
class Matrix
{
public:
  /// Copies the CSR in (Matrix.cpp:13-22) and runs the row-binning analysis.
  Matrix(const CSR<T>& A, std::shared_ptr<L2GMap> col_map, std::shared_ptr<L2GMap> row_map);
#ifdef SPMV_HAVE_EIGEN
  /// The reference's constructor (Matrix.h:30-32).
  Matrix(const Eigen::SparseMatrix<T, Eigen::RowMajor>& A, std::shared_ptr<L2GMap> col_map,
         std::shared_ptr<L2GMap> row_map)
      : Matrix(CSR<T>(A), std::move(col_map), std::move(row_map))
  {
  }
#endif
  ~Matrix();
  Matrix(Matrix&& o) noexcept;
  Matrix(const Matrix&) = delete;
  Matrix& operator=(const Matrix&) = delete;

  /// y = A x (Matrix.h:37-39): x has cols() entries (owned + ghosts).
  Vector<T> operator*(const Vector<T>& b) const;
  /// y = A^T x (Matrix.h:41-43): x has rows() entries.
  Vector<T> transpmult(const Vector<T>& b) const;

  std::shared_ptr<const L2GMap> row_map() const { return _row_map; }
  std::shared_ptr<const L2GMap> col_map() const { return _col_map; }
  std::int64_t rows() const { return _rows; }
  std::int64_t cols() const { return _cols; }
  std::int64_t nonZeros() const { return _nnz; }
  /// Host copy of the stored matrix (Matrix.h:52-53 returns the Eigen
  /// matrix; here it lives in device memory, so this copies it back).
  CSR<T> mat() const;

  spmvx_matrix handle() const { return _h; }

  /// Matrix::create_matrix (Matrix.h:55-65, Matrix.cpp:213-438): the local
  /// block `mat` (nrows_local + row_ghosts.size() rows, ncols_local +
  /// col_ghosts.size() columns, ghost columns numbered by their position in
  /// col_ghosts) becomes a matrix with column ghosts only: ghost columns
  /// renumbered in ascending global order (duplicates merged), entries of
  /// a row sorted by column with duplicates summed (setFromTriplets).  The
  /// reference sends ghost rows to their owners; one process drives one rank
  /// here, so row_ghosts must be empty (runtime_error otherwise).  Without a
  /// Decomposition the rank has no neighbours and col_ghosts must be empty
  /// too; with one, its ghost columns are owned by the other (virtual)
  /// ranks it describes.
  static Matrix<T> create_matrix(MPI_Comm comm, const CSR<T>& mat, std::int64_t nrows_local,
                                 std::int64_t ncols_local, std::vector<std::int64_t> row_ghosts,
                                 std::vector<std::int64_t> col_ghosts);
  static Matrix<T> create_matrix(MPI_Comm comm, const CSR<T>& mat, std::int64_t nrows_local,
                                 std::int64_t ncols_local, std::vector<std::int64_t> row_ghosts,
                                 std::vector<std::int64_t> col_ghosts,
                                 const Decomposition& decomp);
#ifdef SPMV_HAVE_EIGEN
  static Matrix<T> create_matrix(MPI_Comm comm, const Eigen::SparseMatrix<T, Eigen::RowMajor> mat,
                                 std::int64_t nrows_local, std::int64_t ncols_local,
                                 std::vector<std::int64_t> row_ghosts,
                                 std::vector<std::int64_t> col_ghosts)
  {
    return create_matrix(comm, CSR<T>(mat), nrows_local, ncols_local, std::move(row_ghosts),
                         std::move(col_ghosts));
  }
#endif

private:
  static Matrix<T> create_local(MPI_Comm comm, const CSR<T>& mat, std::int64_t nrows_local,
                                std::int64_t ncols_local, const std::vector<std::int64_t>& row_ghosts,
                                const std::vector<std::int64_t>& col_ghosts,
                                const Decomposition* decomp);
  spmvx_matrix _h = nullptr;
  std::int64_t _rows = 0, _cols = 0, _nnz = 0;
  std::shared_ptr<L2GMap> _col_map, _row_map;
};

} // namespace spmv
