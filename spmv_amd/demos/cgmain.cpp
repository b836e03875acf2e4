// CG demo through the C++ API (the structure of demos/cgmain.cpp:16-108).
//   cgmain petsc <name>
//     as the reference: A from petsc_mat<name>.dat, b from petsc_vec<name>.dat
//     (PETSc binary files), x = spmv::cg(A, b, 10000, 1e-10), then the true
//     residual ||A x - b|| and ||x|| (cgmain.cpp:36-81).
//   cgmain [n=512] [rtol=1e-8] [max_its=10000] [seed=1]
//     generated input: 2D 5-point Poisson on an n x n grid (Dirichlet),
//     x_true uniform[-1,1] from a splitmix64 stream, b = A x_true.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include <spmv/spmv.h>

static double uniform_pm1(std::uint64_t seed, std::uint64_t i)
{
  std::uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return -1.0 + 2.0 * (static_cast<double>(z >> 11) * 0x1.0p-53);
}

static int petsc_main(const std::string& name)
{
  auto t0 = std::chrono::steady_clock::now();
  auto A = spmv::read_petsc_binary(MPI_COMM_WORLD, "petsc_mat" + name + ".dat");
  std::shared_ptr<const spmv::L2GMap> l2g = A.col_map();
  spmv::Vector<double> b(spmv::read_petsc_binary_vector(MPI_COMM_WORLD, "petsc_vec" + name + ".dat"));
  std::printf("Global vec size = %lld\n", static_cast<long long>(l2g->global_size()));
  const double read = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  t0 = std::chrono::steady_clock::now();
  auto [x, num_its] = spmv::cg(MPI_COMM_WORLD, A, b, 10000, 1e-10);
  const double solve = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  l2g->update(x.data());
  std::vector<double> ax = (A * x).to_host(), bh = b.to_host(), xh = x.to_host();
  double rn = 0, xn = 0;
  for (std::size_t i = 0; i < bh.size(); ++i)
  {
    rn += (ax[i] - bh[i]) * (ax[i] - bh[i]);
    xn += xh[i] * xh[i];
  }
  std::printf("r.norm = %.17g\n", std::sqrt(rn));
  std::printf("x.norm = %.17g in %d iterations\n", std::sqrt(xn), num_its);
  std::printf("[0.ReadPetsc] %.6f\n[1.Solve] %.6f\n", read, solve);
  return 0;
}

int main(int argc, char** argv)
{
  MPI_Init(&argc, &argv);
  if (argc > 2 && std::string(argv[1]) == "petsc")
  {
    const int rc = petsc_main(argv[2]);
    MPI_Finalize();
    return rc;
  }
  const std::int64_t n = argc > 1 ? std::atoll(argv[1]) : 512;
  const double rtol = argc > 2 ? std::atof(argv[2]) : 1e-8;
  const int max_its = argc > 3 ? std::atoi(argv[3]) : 10000;
  const std::uint64_t seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1;
  const std::int64_t N = n * n;
  spmv::CSR<double> A;
  A.rows = A.cols = N;
  A.outer.push_back(0);
  for (std::int64_t j = 0; j < n; ++j)
    for (std::int64_t i = 0; i < n; ++i)
    {
      const std::int64_t r = i + n * j;
      if (j > 0) A.inner.push_back(static_cast<std::int32_t>(r - n)), A.values.push_back(-1.0);
      if (i > 0) A.inner.push_back(static_cast<std::int32_t>(r - 1)), A.values.push_back(-1.0);
      A.inner.push_back(static_cast<std::int32_t>(r)), A.values.push_back(4.0);
      if (i < n - 1) A.inner.push_back(static_cast<std::int32_t>(r + 1)), A.values.push_back(-1.0);
      if (j < n - 1) A.inner.push_back(static_cast<std::int32_t>(r + n)), A.values.push_back(-1.0);
      A.outer.push_back(static_cast<std::int32_t>(A.inner.size()));
    }
  auto l2g = std::make_shared<spmv::L2GMap>(MPI_COMM_WORLD, N, std::vector<std::int64_t>{});
  auto row = std::make_shared<spmv::L2GMap>(MPI_COMM_WORLD, N, std::vector<std::int64_t>{});
  spmv::Matrix<double> M(A, l2g, row);
  std::vector<double> xt(static_cast<std::size_t>(N));
  for (std::int64_t i = 0; i < N; ++i)
    xt[i] = uniform_pm1(seed, static_cast<std::uint64_t>(i));
  spmv::Vector<double> b = M * spmv::Vector<double>(xt);

  auto t0 = std::chrono::steady_clock::now();
  auto [x, num_its] = spmv::cg(MPI_COMM_WORLD, M, b, max_its, rtol);
  const double solve = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  l2g->update(x.data());
  std::vector<double> ax = (M * x).to_host(), bh = b.to_host(), xh = x.to_host();
  double rn = 0, bn = 0, xn = 0, en = 0;
  for (std::int64_t i = 0; i < N; ++i)
  {
    rn += (ax[i] - bh[i]) * (ax[i] - bh[i]);
    bn += bh[i] * bh[i];
    xn += xh[i] * xh[i];
    en += (xh[i] - xt[i]) * (xh[i] - xt[i]);
  }
  std::printf("iterations = %d\n", num_its);
  std::printf("r.norm = %.17g\n", std::sqrt(rn));
  std::printf("rel.residual = %.6e\n", std::sqrt(rn / bn));
  std::printf("x.norm = %.17g\n", std::sqrt(xn));
  std::printf("rel.error = %.6e\n", std::sqrt(en / xn));
  std::printf("[1.Solve] %.6f\n", solve);
  MPI_Finalize();
  return 0;
}
