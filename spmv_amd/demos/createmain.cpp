// Driver of spmv::Matrix<double>::create_matrix (Matrix.h:55-65) for the
// tests: rank 0 of a two-rank decomposition (rows and columns 0-3 here,
// 4-7 on the other rank) whose local block has unsorted and repeated
// columns, and ghost columns listed out of order and twice.  Prints the
// resulting CSR, the column map's ghosts and y = A x for x = 1, 2, ...;
// then the error of a call with ghost rows.
#include <cstdio>
#include <stdexcept>
#include <vector>

#include <spmv/spmv.h>

int main(int argc, char** argv)
{
  MPI_Init(&argc, &argv);
  int rc = 0;
  try
  {
    spmv::Decomposition d;
    d.local_sizes = {4, 4};
    d.rank = 0;
    d.remote_ghosts = {{}, {3}}; // the other rank ghosts our column 3
    // local columns 0-3 owned, 4-6 ghosts: global 7, 5, 7
    const std::vector<std::int64_t> col_ghosts = {7, 5, 7};
    spmv::CSR<double> A;
    A.rows = 4;
    A.cols = 7;
    A.outer = {0, 4, 6, 9, 11};
    A.inner = {1, 0, 6, 4, /**/ 1, 1, /**/ 5, 2, 3, /**/ 3, 2};
    A.values = {2.0, 4.0, -1.0, -0.5, /**/ 3.0, 0.25, /**/ 7.0, -2.0, 1.5, /**/ 6.0, -3.0};
    spmv::Matrix<double> M =
        spmv::Matrix<double>::create_matrix(MPI_COMM_WORLD, A, 4, 4, {}, col_ghosts, d);
    const spmv::CSR<double> B = M.mat();
    std::printf("rows = %lld\ncols = %lld\nnnz = %lld\n", static_cast<long long>(B.rows),
                static_cast<long long>(B.cols), static_cast<long long>(B.nonZeros()));
    std::printf("outer =");
    for (auto v : B.outer)
      std::printf(" %d", v);
    std::printf("\ninner =");
    for (auto v : B.inner)
      std::printf(" %d", v);
    std::printf("\nvalues =");
    for (auto v : B.values)
      std::printf(" %.17g", v);
    std::printf("\nghosts =");
    for (auto g : M.col_map()->ghosts())
      std::printf(" %lld", static_cast<long long>(g));
    std::vector<double> xh(static_cast<std::size_t>(M.cols()));
    for (std::size_t i = 0; i < xh.size(); ++i)
      xh[i] = static_cast<double>(i + 1);
    const spmv::Vector<double> y = M * spmv::Vector<double>(xh);
    std::printf("\ny =");
    for (double v : y.to_host())
      std::printf(" %.17g", v);
    std::printf("\n");
    try
    {
      spmv::CSR<double> A5 = A;
      A5.rows = 5;
      A5.outer.push_back(A5.outer.back());
      (void)spmv::Matrix<double>::create_matrix(MPI_COMM_WORLD, A5, 4, 4, {5}, col_ghosts, d);
      std::printf("error = none\n");
    }
    catch (const std::runtime_error& e)
    {
      std::printf("error = %s\n", e.what());
    }
  }
  catch (const std::exception& e)
  {
    std::fprintf(stderr, "createmain: %s\n", e.what());
    rc = 1;
  }
  MPI_Finalize();
  return rc;
}
