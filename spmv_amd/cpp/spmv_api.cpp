// C++ API (spmv::Vector, spmv::L2GMap, spmv::Matrix, spmv::cg) over the
// spmvx C ABI.  Errors become std::runtime_error carrying the library's
// message, which for the reference's checks is the reference's text.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "spmv/spmv.h"

namespace spmv
{
namespace
{

void check(int status)
{
  if (status != SPMVX_OK)
    throw std::runtime_error(spmvx_last_error());
}

void hip_check(hipError_t e, const char* what)
{
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
constexpr int32_t dtype_of()
{
  if (std::is_same<T, double>::value)
    return SPMVX_F64;
  if (std::is_same<T, std::complex<double>>::value)
    return SPMVX_C64;
  if (std::is_same<T, std::complex<float>>::value)
    return SPMVX_C32;
  return SPMVX_F32;
}

// Host values of a CSR as the doubles the C ABI takes ((re, im) pairs for
// complex values).
template <typename T>
std::vector<double> as_doubles(const std::vector<T>& v)
{
  return std::vector<double>(v.begin(), v.end());
}
template <typename R>
std::vector<double> as_doubles(const std::vector<std::complex<R>>& v)
{
  std::vector<double> d(2 * v.size());
  for (std::size_t i = 0; i < v.size(); ++i)
  {
    d[2 * i] = static_cast<double>(v[i].real());
    d[2 * i + 1] = static_cast<double>(v[i].imag());
  }
  return d;
}

int comm_size(MPI_Comm comm)
{
  int size = 1;
  MPI_Comm_size(comm, &size);
  return size;
}

} // namespace

// ---------------------------------------------------------------- Vector
template <typename T>
Vector<T>::Vector(std::size_t n) : _n(n)
{
  if (n)
    hip_check(hipMalloc(&_p, n * sizeof(T)), "hipMalloc");
  setZero();
}

template <typename T>
Vector<T>::Vector(const std::vector<T>& host) : Vector(host.size())
{
  from_host(host);
}

template <typename T>
Vector<T>::~Vector()
{
  if (_p)
    (void)hipFree(_p);
}

template <typename T>
Vector<T>::Vector(Vector&& o) noexcept : _p(o._p), _n(o._n)
{
  o._p = nullptr;
  o._n = 0;
}

template <typename T>
Vector<T>& Vector<T>::operator=(Vector&& o) noexcept
{
  if (this != &o)
  {
    if (_p)
      (void)hipFree(_p);
    _p = o._p;
    _n = o._n;
    o._p = nullptr;
    o._n = 0;
  }
  return *this;
}

template <typename T>
Vector<T> Vector<T>::clone() const
{
  Vector<T> v(_n);
  if (_n)
    hip_check(hipMemcpy(v._p, _p, _n * sizeof(T), hipMemcpyDeviceToDevice), "hipMemcpy");
  return v;
}

template <typename T>
void Vector<T>::setZero()
{
  if (_n)
    hip_check(hipMemset(_p, 0, _n * sizeof(T)), "hipMemset");
}

template <typename T>
void Vector<T>::copy_head(const Vector& src, std::size_t n)
{
  if (n > _n || n > src._n)
    throw std::runtime_error("copy_head: vectors too short");
  if (n)
    hip_check(hipMemcpy(_p, src._p, n * sizeof(T), hipMemcpyDeviceToDevice), "hipMemcpy");
}

template <typename T>
std::vector<T> Vector<T>::to_host() const
{
  std::vector<T> h(_n);
  if (_n)
    hip_check(hipMemcpy(h.data(), _p, _n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  return h;
}

template <typename T>
void Vector<T>::from_host(const std::vector<T>& host)
{
  if (host.size() != _n)
    throw std::runtime_error("from_host: size mismatch");
  if (_n)
    hip_check(hipMemcpy(_p, host.data(), _n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
}

template class Vector<double>;
template class Vector<float>;
template class Vector<std::complex<double>>;
template class Vector<std::complex<float>>;

// ---------------------------------------------------------------- L2GMap
L2GMap::L2GMap(MPI_Comm comm, std::int64_t local_size, const std::vector<std::int64_t>& ghosts)
    : _ghosts(ghosts)
{
  Decomposition d;
  d.local_sizes = {local_size};
  d.rank = 0;
  build(comm, local_size, d);
}

L2GMap::L2GMap(MPI_Comm comm, std::int64_t local_size, const std::vector<std::int64_t>& ghosts,
               const Decomposition& decomp)
    : _ghosts(ghosts)
{
  build(comm, local_size, decomp);
}

void L2GMap::build(MPI_Comm comm, std::int64_t local_size, const Decomposition& d)
{
  if (comm_size(comm) != 1)
    throw std::runtime_error("spmv::L2GMap: one process drives one rank here; a communicator "
                             "with several processes needs the multi-GPU transport");
  const int nranks = static_cast<int>(d.local_sizes.size());
  if (nranks < 1 || d.rank < 0 || d.rank >= nranks || d.local_sizes[d.rank] != local_size)
    throw std::runtime_error("spmv::L2GMap: inconsistent decomposition");
  std::vector<std::int64_t> rptr(nranks + 1, 0), ridx;
  for (int q = 0; q < nranks; ++q)
  {
    if (q < static_cast<int>(d.remote_ghosts.size()))
      ridx.insert(ridx.end(), d.remote_ghosts[q].begin(), d.remote_ghosts[q].end());
    rptr[q + 1] = static_cast<std::int64_t>(ridx.size());
  }
  check(spmvx_l2gmap_create(nranks, d.rank, d.local_sizes.data(), _ghosts.data(),
                            static_cast<std::int64_t>(_ghosts.size()), rptr.data(),
                            ridx.empty() ? nullptr : ridx.data(), &_h));
}

L2GMap::~L2GMap()
{
  if (_h)
    spmvx_l2gmap_destroy(_h);
}

std::int32_t L2GMap::local_size() const
{
  std::int64_t v = 0;
  check(spmvx_l2gmap_info(_h, &v, nullptr, nullptr, nullptr, nullptr, nullptr));
  return static_cast<std::int32_t>(v);
}

std::int32_t L2GMap::num_ghosts() const { return static_cast<std::int32_t>(_ghosts.size()); }

std::int64_t L2GMap::global_size() const
{
  std::int64_t v = 0;
  check(spmvx_l2gmap_info(_h, nullptr, nullptr, &v, nullptr, nullptr, nullptr));
  return v;
}

std::int64_t L2GMap::global_offset() const
{
  std::int64_t v = 0;
  check(spmvx_l2gmap_info(_h, nullptr, nullptr, nullptr, &v, nullptr, nullptr));
  return v;
}

std::int32_t L2GMap::global_to_local(std::int64_t i) const
{
  std::int32_t v = 0;
  check(spmvx_l2gmap_global_to_local(_h, i, &v));
  return v;
}

template <typename T>
void L2GMap::update(T* vec_data) const
{
  check(spmvx_l2gmap_update(_h, dtype_of<T>(), vec_data, nullptr, nullptr));
}

template <typename T>
void L2GMap::reverse_update(T* vec_data, const std::vector<T>& incoming) const
{
  std::int64_t nidx = 0;
  check(spmvx_l2gmap_info(_h, nullptr, nullptr, nullptr, nullptr, &nidx, nullptr));
  if (!incoming.empty() && static_cast<std::int64_t>(incoming.size()) != nidx)
    throw std::runtime_error("reverse_update: incoming has the wrong length");
  check(spmvx_l2gmap_reverse_update(_h, dtype_of<T>(), vec_data,
                                    incoming.empty() ? nullptr : incoming.data(), nullptr,
                                    nullptr));
}

template <typename T>
void L2GMap::set_halo(const std::vector<T>& halo)
{
  if (halo.size() != _ghosts.size())
    throw std::runtime_error("set_halo: length differs from num_ghosts");
  check(spmvx_l2gmap_set_halo(_h, dtype_of<T>(), halo.data()));
}

void L2GMap::clear_halo() { check(spmvx_l2gmap_set_halo(_h, SPMVX_F64, nullptr)); }

template void L2GMap::update<double>(double*) const;
template void L2GMap::update<float>(float*) const;
template void L2GMap::reverse_update<double>(double*, const std::vector<double>&) const;
template void L2GMap::reverse_update<float>(float*, const std::vector<float>&) const;
template void L2GMap::set_halo<double>(const std::vector<double>&);
template void L2GMap::set_halo<float>(const std::vector<float>&);
// complex instantiations (L2GMap.cpp:202-213)
template void L2GMap::update<std::complex<double>>(std::complex<double>*) const;
template void L2GMap::update<std::complex<float>>(std::complex<float>*) const;
template void L2GMap::reverse_update<std::complex<double>>(
    std::complex<double>*, const std::vector<std::complex<double>>&) const;
template void L2GMap::reverse_update<std::complex<float>>(
    std::complex<float>*, const std::vector<std::complex<float>>&) const;
template void L2GMap::set_halo<std::complex<double>>(const std::vector<std::complex<double>>&);
template void L2GMap::set_halo<std::complex<float>>(const std::vector<std::complex<float>>&);

// ---------------------------------------------------------------- Matrix
template <typename T>
Matrix<T>::Matrix(const CSR<T>& A, std::shared_ptr<L2GMap> col_map,
                  std::shared_ptr<L2GMap> row_map)
    : _rows(A.rows), _cols(A.cols), _nnz(A.nonZeros()), _col_map(std::move(col_map)),
      _row_map(std::move(row_map))
{
  if (static_cast<std::int64_t>(A.outer.size()) != A.rows + 1
      || static_cast<std::int64_t>(A.inner.size()) != _nnz
      || static_cast<std::int64_t>(A.values.size()) != _nnz)
    throw std::runtime_error("spmv::Matrix: inconsistent CSR arrays");
  if (_col_map && _col_map->local_size() + _col_map->num_ghosts() != A.cols)
    throw std::runtime_error("spmv::Matrix: column map size differs from the column count");
  const std::vector<double> v = as_doubles(A.values);
  check(spmvx_matrix_create(dtype_of<T>(), A.rows, A.cols, _nnz, A.outer.data(), A.inner.data(),
                            v.data(), &_h));
}

template <typename T>
Matrix<T>::~Matrix()
{
  if (_h)
    spmvx_matrix_destroy(_h);
}

template <typename T>
Matrix<T>::Matrix(Matrix&& o) noexcept
    : _h(o._h), _rows(o._rows), _cols(o._cols), _nnz(o._nnz), _col_map(std::move(o._col_map)),
      _row_map(std::move(o._row_map))
{
  o._h = nullptr;
}

template <typename T>
Vector<T> Matrix<T>::operator*(const Vector<T>& b) const
{
  if (static_cast<std::int64_t>(b.size()) != _cols)
    throw std::runtime_error("spmv::Matrix::operator*: vector size differs from cols()");
  Vector<T> y(static_cast<std::size_t>(_rows));
  check(spmvx_spmv(_h, b.data(), y.data(), nullptr));
  return y;
}

template <typename T>
Vector<T> Matrix<T>::transpmult(const Vector<T>& b) const
{
  if (static_cast<std::int64_t>(b.size()) != _rows)
    throw std::runtime_error("spmv::Matrix::transpmult: vector size differs from rows()");
  Vector<T> y(static_cast<std::size_t>(_cols));
  check(spmvx_spmv_t(_h, b.data(), y.data(), nullptr));
  return y;
}

template <typename T>
CSR<T> Matrix<T>::mat() const
{
  CSR<T> A;
  A.rows = _rows;
  A.cols = _cols;
  A.outer.resize(static_cast<std::size_t>(_rows) + 1);
  A.inner.resize(static_cast<std::size_t>(_nnz));
  A.values.resize(static_cast<std::size_t>(_nnz));
  const std::int32_t *rp = nullptr, *ci = nullptr;
  const void* va = nullptr;
  check(spmvx_matrix_arrays(_h, &rp, &ci, &va));
  hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  hip_check(hipMemcpy(A.outer.data(), rp, sizeof(std::int32_t) * A.outer.size(),
                      hipMemcpyDeviceToHost), "hipMemcpy");
  if (_nnz > 0)
  {
    hip_check(hipMemcpy(A.inner.data(), ci, sizeof(std::int32_t) * A.inner.size(),
                        hipMemcpyDeviceToHost), "hipMemcpy");
    // the stored values have T's layout (F64, F32, C64, C32)
    hip_check(hipMemcpy(A.values.data(), va, sizeof(T) * A.values.size(), hipMemcpyDeviceToHost),
              "hipMemcpy");
  }
  return A;
}

template <typename T>
Matrix<T> Matrix<T>::create_local(MPI_Comm comm, const CSR<T>& mat, std::int64_t nrows_local,
                                  std::int64_t ncols_local,
                                  const std::vector<std::int64_t>& row_ghosts,
                                  const std::vector<std::int64_t>& col_ghosts,
                                  const Decomposition* decomp)
{
  if (!row_ghosts.empty())
    throw std::runtime_error("spmv::Matrix::create_matrix: ghost rows go to the ranks that own "
                             "them; one process drives one rank, so row_ghosts must be empty");
  if (nrows_local < 0 || ncols_local < 0 || mat.rows != nrows_local
      || mat.cols != ncols_local + static_cast<std::int64_t>(col_ghosts.size())
      || static_cast<std::int64_t>(mat.outer.size()) != mat.rows + 1
      || mat.inner.size() != static_cast<std::size_t>(mat.nonZeros())
      || mat.values.size() != static_cast<std::size_t>(mat.nonZeros()))
    throw std::runtime_error("spmv::Matrix::create_matrix: matrix sizes differ from the maps");
  // ghost columns in ascending global order, each once (Matrix.cpp:342-367)
  std::vector<std::int64_t> ghosts(col_ghosts);
  std::sort(ghosts.begin(), ghosts.end());
  ghosts.erase(std::unique(ghosts.begin(), ghosts.end()), ghosts.end());
  auto local_col = [&](std::int32_t c) -> std::int32_t {
    if (c < 0 || c >= mat.cols)
      throw std::runtime_error("spmv::Matrix::create_matrix: column index out of range");
    if (c < ncols_local)
      return c;
    const std::int64_t g = col_ghosts[static_cast<std::size_t>(c - ncols_local)];
    return static_cast<std::int32_t>(
        ncols_local + (std::lower_bound(ghosts.begin(), ghosts.end(), g) - ghosts.begin()));
  };
  // rows: entries sorted by column, duplicates summed in stored order
  // (Eigen's setFromTriplets, Matrix.cpp:370-384,421-424)
  CSR<T> B;
  B.rows = nrows_local;
  B.cols = ncols_local + static_cast<std::int64_t>(ghosts.size());
  B.outer.assign(static_cast<std::size_t>(nrows_local) + 1, 0);
  std::vector<std::pair<std::int32_t, T>> row;
  for (std::int64_t r = 0; r < nrows_local; ++r)
  {
    row.clear();
    for (std::int32_t k = mat.outer[static_cast<std::size_t>(r)];
         k < mat.outer[static_cast<std::size_t>(r) + 1]; ++k)
      row.emplace_back(local_col(mat.inner[static_cast<std::size_t>(k)]),
                       mat.values[static_cast<std::size_t>(k)]);
    std::stable_sort(row.begin(), row.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    for (std::size_t i = 0; i < row.size(); ++i)
    {
      if (i > 0 && row[i].first == row[i - 1].first)
        B.values.back() += row[i].second;
      else
      {
        B.inner.push_back(row[i].first);
        B.values.push_back(row[i].second);
      }
    }
    B.outer[static_cast<std::size_t>(r) + 1] = static_cast<std::int32_t>(B.inner.size());
  }
  std::shared_ptr<L2GMap> col_map, row_map;
  if (decomp)
  {
    Decomposition rows_only = *decomp;
    for (auto& g : rows_only.remote_ghosts)
      g.clear();
    col_map = std::make_shared<L2GMap>(comm, ncols_local, ghosts, *decomp);
    row_map = std::make_shared<L2GMap>(comm, nrows_local, std::vector<std::int64_t>(), rows_only);
  }
  else
  {
    col_map = std::make_shared<L2GMap>(comm, ncols_local, ghosts);
    row_map = std::make_shared<L2GMap>(comm, nrows_local, std::vector<std::int64_t>());
  }
  return Matrix<T>(B, col_map, row_map);
}

template <typename T>
Matrix<T> Matrix<T>::create_matrix(MPI_Comm comm, const CSR<T>& mat, std::int64_t nrows_local,
                                   std::int64_t ncols_local, std::vector<std::int64_t> row_ghosts,
                                   std::vector<std::int64_t> col_ghosts)
{
  return create_local(comm, mat, nrows_local, ncols_local, row_ghosts, col_ghosts, nullptr);
}

template <typename T>
Matrix<T> Matrix<T>::create_matrix(MPI_Comm comm, const CSR<T>& mat, std::int64_t nrows_local,
                                   std::int64_t ncols_local, std::vector<std::int64_t> row_ghosts,
                                   std::vector<std::int64_t> col_ghosts,
                                   const Decomposition& decomp)
{
  return create_local(comm, mat, nrows_local, ncols_local, row_ghosts, col_ghosts, &decomp);
}

template class Matrix<double>;
template class Matrix<float>;
template class Matrix<std::complex<double>>; // Matrix.cpp:441-444
template class Matrix<std::complex<float>>;

// ---------------------------------------------------------------- cg
std::tuple<Vector<double>, int> cg(MPI_Comm comm, const Matrix<double>& A,
                                   const Vector<double>& b, int max_its, double rtol)
{
  (void)comm_size(comm);
  std::shared_ptr<const L2GMap> col = A.col_map(), row = A.row_map();
  if (row && row->num_ghosts() > 0) // cg.cpp:21-23
    throw std::runtime_error("spmv::cg - Error: A.row_map() has ghost entries");
  const std::int64_t M = row ? row->local_size() : A.rows();
  if (static_cast<std::int64_t>(b.size()) != M) // cg.cpp:27-28
    throw std::runtime_error("spmv::cg - Error: b.rows() != A.rows()");
  Vector<double> x(static_cast<std::size_t>(A.cols()));
  std::int32_t its = 0;
  check(spmvx_cg(A.handle(), col ? col->handle() : nullptr, b.data(), x.data(), max_its, rtol,
                 &its, nullptr, nullptr));
  return std::make_tuple(std::move(x), static_cast<int>(its));
}

// ---------------------------------------------------------------- PETSc input
namespace
{
std::vector<std::int64_t> split(int size, std::int64_t N) // read_petsc.cpp:13-30
{
  const std::int64_t n = N / size, r = N % size;
  std::vector<std::int64_t> sizes(size);
  for (int q = 0; q < size; ++q)
    sizes[q] = n + (q < r ? 1 : 0);
  return sizes;
}
} // namespace

Matrix<double> read_petsc_binary(MPI_Comm comm, std::string filename, int nranks, int rank)
{
  if (comm_size(comm) != 1)
    throw std::runtime_error("spmv::read_petsc_binary: one process drives one rank here");
  spmvx_petsc_csr h = nullptr;
  check(spmvx_petsc_csr_read(filename.c_str(), nranks, rank, &h));
  std::unique_ptr<spmvx_petsc_csr_s, int (*)(spmvx_petsc_csr)> hold(h, spmvx_petsc_csr_destroy);
  std::int64_t nr = 0, nc = 0, nrl = 0, ncl = 0, ng = 0, nnz = 0;
  check(spmvx_petsc_csr_info(h, &nr, &nc, &nrl, &ncl, &ng, &nnz));
  if (rank == 0)
    std::printf("Read file: %s: %lldx%lld = %lld\n", filename.c_str(),
                static_cast<long long>(nr), static_cast<long long>(nc),
                static_cast<long long>(nnz)); // read_petsc.cpp:72-74
  const std::int32_t *rp = nullptr, *ci = nullptr;
  const double* val = nullptr;
  const std::int64_t* gh = nullptr;
  check(spmvx_petsc_csr_arrays(h, &rp, &ci, &val, &gh));
  const std::int64_t *rptr = nullptr, *rgh = nullptr;
  check(spmvx_petsc_csr_remote(h, &rptr, &rgh));
  CSR<double> A;
  A.rows = nrl;
  A.cols = ncl + ng;
  A.outer.assign(rp, rp + nrl + 1);
  A.inner.assign(ci, ci + nnz);
  A.values.assign(val, val + nnz);
  Decomposition dc, dr;
  dc.local_sizes = split(nranks, nc);
  dc.rank = rank;
  dc.remote_ghosts.resize(nranks);
  for (int q = 0; q < nranks; ++q)
    dc.remote_ghosts[q].assign(rgh + rptr[q], rgh + rptr[q + 1]);
  dr.local_sizes = split(nranks, nr);
  dr.rank = rank;
  auto col_map = std::make_shared<L2GMap>(comm, ncl, std::vector<std::int64_t>(gh, gh + ng), dc);
  auto row_map = std::make_shared<L2GMap>(comm, nrl, std::vector<std::int64_t>{}, dr);
  return Matrix<double>(A, col_map, row_map);
}

Matrix<double> read_petsc_binary(MPI_Comm comm, std::string filename)
{
  return read_petsc_binary(comm, std::move(filename), 1, 0);
}

std::vector<double> read_petsc_binary_vector(MPI_Comm comm, std::string filename, int nranks,
                                             int rank)
{
  if (comm_size(comm) != 1)
    throw std::runtime_error("spmv::read_petsc_binary_vector: one process drives one rank here");
  std::int64_t nl = 0, ng = 0;
  check(spmvx_petsc_read_vector(filename.c_str(), nranks, rank, nullptr, 0, &nl, &ng));
  if (rank == 0)
    std::printf("Read vector file: %s: %lld\n", filename.c_str(),
                static_cast<long long>(ng)); // read_petsc.cpp:228-229
  std::vector<double> v(static_cast<std::size_t>(nl));
  check(spmvx_petsc_read_vector(filename.c_str(), nranks, rank, v.data(), nl, nullptr, nullptr));
  return v;
}

std::vector<double> read_petsc_binary_vector(MPI_Comm comm, std::string filename)
{
  return read_petsc_binary_vector(comm, std::move(filename), 1, 0);
}

} // namespace spmv
