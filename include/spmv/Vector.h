// Device vectors and host CSR containers of the C++ API.  They take the place
// of the Eigen types in the reference's signatures (Eigen::VectorXd /
// Eigen::SparseMatrix<T, RowMajor>, spmv/Matrix.h:30-43): vectors live in
// device memory (HBM); the CSR is a plain compressed row-major host matrix
// with the same three arrays Eigen exposes (outerIndexPtr, innerIndexPtr,
// valuePtr).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

// Eigen adaptor: where Eigen is installed, the containers also convert from
// and to the reference's Eigen types (plain element copies).
#if defined(__has_include)
#if __has_include(<Eigen/Sparse>)
#include <Eigen/Dense>
#include <Eigen/Sparse>
#define SPMV_HAVE_EIGEN 1
#endif
#endif

namespace spmv
{

/// Compressed row-major sparse matrix on the host (rows x cols), column
/// indices sorted within each row.
template <typename T>
struct CSR
{
  std::int64_t rows = 0, cols = 0;
  std::vector<std::int32_t> outer; ///< rows + 1 row offsets
  std::vector<std::int32_t> inner; ///< column index per entry
  std::vector<T> values;           ///< value per entry
  std::int64_t nonZeros() const { return outer.empty() ? 0 : outer.back(); }

#ifdef SPMV_HAVE_EIGEN
  CSR() = default;
  /// From the reference's input type (Matrix.h:30-32): the entries of every
  /// row in stored order (compressed or not).
  explicit CSR(const Eigen::SparseMatrix<T, Eigen::RowMajor>& A) : rows(A.rows()), cols(A.cols())
  {
    outer.assign(static_cast<std::size_t>(rows) + 1, 0);
    for (Eigen::Index r = 0; r < A.outerSize(); ++r)
    {
      for (typename Eigen::SparseMatrix<T, Eigen::RowMajor>::InnerIterator it(A, r); it; ++it)
      {
        inner.push_back(static_cast<std::int32_t>(it.col()));
        values.push_back(it.value());
      }
      outer[static_cast<std::size_t>(r) + 1] = static_cast<std::int32_t>(inner.size());
    }
  }
  /// The same matrix as Eigen stores it (what Matrix::mat() returns in the
  /// reference, Matrix.h:52-53).
  Eigen::SparseMatrix<T, Eigen::RowMajor> to_eigen() const
  {
    return Eigen::Map<const Eigen::SparseMatrix<T, Eigen::RowMajor>>(
        rows, cols, nonZeros(), outer.data(), inner.data(), values.data());
  }
#endif
};

/// Vector in device memory (owning, movable, not copyable implicitly).
template <typename T>
class Vector
{
public:
  Vector() = default;
  explicit Vector(std::size_t n);
  explicit Vector(const std::vector<T>& host);
  ~Vector();
  Vector(Vector&& o) noexcept;
  Vector& operator=(Vector&& o) noexcept;
  Vector(const Vector&) = delete;
  Vector& operator=(const Vector&) = delete;
  /// Deep copy on the device.
  Vector clone() const;
#ifdef SPMV_HAVE_EIGEN
  /// From / to the reference's vector type (Eigen::VectorXd for T = double).
  explicit Vector(const Eigen::Matrix<T, Eigen::Dynamic, 1>& host)
      : Vector(std::vector<T>(host.data(), host.data() + host.size()))
  {
  }
  Eigen::Matrix<T, Eigen::Dynamic, 1> to_eigen() const
  {
    const std::vector<T> h = to_host();
    return Eigen::Map<const Eigen::Matrix<T, Eigen::Dynamic, 1>>(
        h.data(), static_cast<Eigen::Index>(h.size()));
  }
#endif

  std::size_t size() const { return _n; }
  std::size_t rows() const { return _n; }
  T* data() { return _p; }
  const T* data() const { return _p; }

  void setZero();
  /// Copies the first n entries of `src` (device) into the first n entries.
  void copy_head(const Vector& src, std::size_t n);
  std::vector<T> to_host() const;
  void from_host(const std::vector<T>& host);

private:
  T* _p = nullptr;
  std::size_t _n = 0;
};

} // namespace spmv
