// Restriction / prolongation demo through the C++ API (the structure of
// demos/restrictmain.cpp:15-110): R from a PETSc binary matrix file, q from a
// PETSc binary vector file, then 10 times
//   psp = R^T q; col_map->reverse_update(psp); ||psp||^2 (owned part)
//   col_map->update(psp); q = R psp; ||q||^2
//   restrictmain [R.dat=R4.dat] [b.dat=b4.dat] [complex]
// With "complex" the values become R (1 + 0.5i) and q (1 - 0.25i) and the
// products run on Matrix<std::complex<double>>.
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <spmv/spmv.h>

template <typename T>
static double sqnorm(const std::vector<T>& v, std::size_t n)
{
  double s = 0;
  for (std::size_t i = 0; i < n; ++i)
    s += std::norm(std::complex<double>(v[i]));
  return s;
}

template <typename T>
static int run(const std::string& rfile, const std::string& bfile, T rscale, T qscale)
{
  auto t0 = std::chrono::steady_clock::now();
  spmv::Matrix<double> R0 = spmv::read_petsc_binary(MPI_COMM_WORLD, rfile);
  std::vector<double> q0 = spmv::read_petsc_binary_vector(MPI_COMM_WORLD, bfile);
  auto l2g = R0.col_map();
  std::printf("Vector = %zu %lld\n", q0.size(), static_cast<long long>(R0.rows()));
  // the value type of the run: the file's matrix rebuilt with scaled values
  spmv::CSR<T> A;
  A.rows = R0.rows();
  A.cols = R0.cols();
  {
    // the same local CSR, from the host-side parse of the C ABI
    spmvx_petsc_csr h = nullptr;
    if (spmvx_petsc_csr_read(rfile.c_str(), 1, 0, &h) != SPMVX_OK)
      throw std::runtime_error(spmvx_last_error());
    const std::int32_t *rp = nullptr, *ci = nullptr;
    const double* v = nullptr;
    spmvx_petsc_csr_arrays(h, &rp, &ci, &v, nullptr);
    const std::int64_t nnz = R0.nonZeros();
    A.outer.assign(rp, rp + A.rows + 1);
    A.inner.assign(ci, ci + nnz);
    A.values.resize(static_cast<std::size_t>(nnz));
    for (std::int64_t k = 0; k < nnz; ++k)
      A.values[k] = T(v[k]) * rscale;
    spmvx_petsc_csr_destroy(h);
  }
  auto cmap = std::const_pointer_cast<spmv::L2GMap>(l2g);
  auto rmap = std::const_pointer_cast<spmv::L2GMap>(R0.row_map());
  spmv::Matrix<T> R(A, cmap, rmap);
  {
    // Matrix::mat() (Matrix.h:52-53): the stored matrix read back equals the
    // one handed in
    const spmv::CSR<T> back = R.mat();
    const bool same = back.rows == A.rows && back.cols == A.cols && back.outer == A.outer
                      && back.inner == A.inner && back.values == A.values;
    std::printf("mat.equal = %d\n", same ? 1 : 0);
  }
  std::vector<T> qh(q0.size());
  for (std::size_t i = 0; i < q0.size(); ++i)
    qh[i] = T(q0[i]) * qscale;
  spmv::Vector<T> q(qh);
  const double read = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  double pnorm = 0, qnorm = 0;
  t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < 10; ++i)
  {
    spmv::Vector<T> psp = R.transpmult(q); // restrict
    l2g->reverse_update(psp.data());
    pnorm = sqnorm(psp.to_host(), static_cast<std::size_t>(l2g->local_size()));
    l2g->update(psp.data()); // prolongate
    q = R * psp;
    qnorm = sqnorm(q.to_host(), static_cast<std::size_t>(q.size()));
  }
  const double loop = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("pnorm = %.17g\nqnorm = %.17g\n", pnorm, qnorm);
  std::printf("[0.PetscRead] %.6f\n[3.SpMV+2.SparseUpdate] %.6f\n", read, loop);
  return 0;
}

int main(int argc, char** argv)
{
  MPI_Init(&argc, &argv);
  const std::string rfile = argc > 1 ? argv[1] : "R4.dat";
  const std::string bfile = argc > 2 ? argv[2] : "b4.dat";
  const bool cplx = argc > 3 && std::strcmp(argv[3], "complex") == 0;
  int rc;
  if (cplx)
    rc = run<std::complex<double>>(rfile, bfile, {1.0, 0.5}, {1.0, -0.25});
  else
    rc = run<double>(rfile, bfile, 1.0, 1.0);
  MPI_Finalize();
  return rc;
}
