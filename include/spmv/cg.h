// spmv::cg -- conjugate gradient solver (the interface of spmv/cg.h:27-29)
// running on the device.
#pragma once

#include <tuple>

#include "Matrix.h"

namespace spmv
{

/// Solves A x = b from x = 0 until ||r||^2/||b||^2 < rtol^2 or max_its
/// (spmv/cg.cpp:11-83).  Returns (x, iterations); x has A.col_map() size with
/// zero ghosts.
std::tuple<Vector<double>, int> cg(MPI_Comm comm, const Matrix<double>& A,
                                   const Vector<double>& b, int max_its, double rtol);

#ifdef SPMV_HAVE_EIGEN
/// The reference's signature (cg.h:27-29): b and x in host memory.
inline std::tuple<Eigen::VectorXd, int> cg(MPI_Comm comm, const Matrix<double>& A,
                                           const Eigen::Ref<const Eigen::VectorXd>& b,
                                           int max_its, double rtol)
{
  const Vector<double> bd(std::vector<double>(b.data(), b.data() + b.size()));
  auto [x, k] = cg(comm, A, bd, max_its, rtol);
  return {x.to_eigen(), k};
}
#endif

} // namespace spmv
