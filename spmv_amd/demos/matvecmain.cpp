// The reference's matvec demo (demos/matvecmain.cpp:16-121) on one rank,
// through the C++ API: 1D 3-point matrix of demos/CreateA.cpp:52-74, vector
// initialised as a Gaussian bump, n_apply x (update, SpMV, copy), norm.
//   matvecmain [N=500000] [n_apply=1000]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>

#include <spmv/spmv.h>

int main(int argc, char** argv)
{
  MPI_Init(&argc, &argv);
  const std::int64_t N = argc > 1 ? std::atoll(argv[1]) : 500000;
  const int n_apply = argc > 2 ? std::atoi(argv[2]) : 1000;
  std::map<std::string, double> timings;
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto t0 = now();

  // create_A: rows (gamma, 1-2 gamma, gamma), first (1-g, g), last (g, 1-g)
  const double gamma = 0.1;
  spmv::CSR<double> A;
  A.rows = A.cols = N;
  A.outer.push_back(0);
  for (std::int64_t c0 = 0; c0 < N; ++c0)
  {
    if (c0 == 0)
    {
      A.inner.push_back(0), A.values.push_back(1.0 - gamma);
      if (N > 1)
        A.inner.push_back(1), A.values.push_back(gamma);
    }
    else if (c0 == N - 1)
    {
      A.inner.push_back(static_cast<std::int32_t>(c0 - 1)), A.values.push_back(gamma);
      A.inner.push_back(static_cast<std::int32_t>(c0)), A.values.push_back(1.0 - gamma);
    }
    else
    {
      A.inner.push_back(static_cast<std::int32_t>(c0 - 1)), A.values.push_back(gamma);
      A.inner.push_back(static_cast<std::int32_t>(c0)), A.values.push_back(1.0 - 2.0 * gamma);
      A.inner.push_back(static_cast<std::int32_t>(c0 + 1)), A.values.push_back(gamma);
    }
    A.outer.push_back(static_cast<std::int32_t>(A.inner.size()));
  }
  auto l2g = std::make_shared<spmv::L2GMap>(MPI_COMM_WORLD, N, std::vector<std::int64_t>{});
  auto row = std::make_shared<spmv::L2GMap>(MPI_COMM_WORLD, N, std::vector<std::int64_t>{});
  spmv::Matrix<double> M(A, l2g, row);
  timings["0.MatCreate"] = std::chrono::duration<double>(now() - t0).count();

  t0 = now();
  std::vector<double> h(static_cast<std::size_t>(l2g->local_size() + l2g->num_ghosts()));
  for (std::int64_t i = 0; i < N; ++i)
  {
    const double z = static_cast<double>(i) / static_cast<double>(N);
    h[i] = std::exp(-10 * std::pow(5 * (z - 0.5), 2.0));
  }
  spmv::Vector<double> psp(h);
  timings["1.VecCreate"] = std::chrono::duration<double>(now() - t0).count();

  t0 = now();
  for (int i = 0; i < n_apply; ++i)
  {
    l2g->update(psp.data());
    spmv::Vector<double> q = M * psp;
    psp.copy_head(q, static_cast<std::size_t>(N));
  }
  std::vector<double> out = psp.to_host();
  timings["2.Apply"] = std::chrono::duration<double>(now() - t0).count();

  double pnorm = 0.0;
  for (std::int64_t i = 0; i < N; ++i)
    pnorm += out[i] * out[i];
  for (auto& t : timings)
    std::printf("[%s] %.6f\n", t.first.c_str(), t.second);
  std::printf("norm = %.17g\n", pnorm);
  MPI_Finalize();
  return 0;
}
