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
