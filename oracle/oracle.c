/*
 * CPU oracle for the libspmv hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C, the arithmetic of the reference library
 * (chrisrichardson/spmv) on the path
 *   Matrix<double>::operator*   spmv/Matrix.cpp:199-204 (Eigen row-major product)
 *   L2GMap::update              spmv/L2GMap.cpp:126-150
 *   L2GMap::reverse_update      spmv/L2GMap.cpp:152-170
 *   spmv::cg                    spmv/cg.cpp:11-83
 * plus the synthetic problem generators the parity tests and the CPU baseline
 * use.  It is only ever called by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker or the timed CPU path,
 * never as part of the GPU product.
 *
 * Rounding contract (what makes the GPU "ordered" SpMV bit-comparable):
 *   Eigen 3.3's sparse(row-major) * dense product evaluates, for every row i,
 *     tmp = 0;  for k in row i (stored order): tmp += a_k * x[col_k];
 *     y_i = 0 + 1 * tmp
 *   with every product and every sum rounded separately (x86-64 without
 *   -march has no FMA; CMakeLists.txt:32).  This file is compiled with
 *   -ffp-contract=off so gcc keeps that rounding.
 *
 * Reductions (dot, squared norm) are sequential left-to-right sums.  Eigen's
 * SIMD reduction order cannot be restated without Eigen; CG parity is therefore
 * defined with a tolerance (iteration count +-1, residual history 1e-10).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* splitmix64 counter-based RNG: value i of stream `seed` is                */
/* mix(seed + (i+1)*gamma).  Identical formula in numpy (tests/bench).       */
/* ------------------------------------------------------------------------ */
static inline uint64_t or_mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline uint64_t or_rand_u64(uint64_t seed, uint64_t i)
{
  return or_mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
static inline double or_rand_u01(uint64_t seed, uint64_t i)
{
  return (double)(or_rand_u64(seed, i) >> 11) * 0x1.0p-53;
}

void or_uniform(uint64_t seed, int64_t n, double lo, double hi, double* out)
{
  for (int64_t i = 0; i < n; ++i)
    out[i] = lo + (hi - lo) * or_rand_u01(seed, (uint64_t)i);
}

/* ------------------------------------------------------------------------ */
/* SpMV: spmv/Matrix.cpp:199-204 (generic path, `_matA * b`)                 */
/* ------------------------------------------------------------------------ */
void or_spmv_f64(int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                 const double* val, const double* x, double* y)
{
  for (int64_t i = 0; i < nrows; ++i)
  {
    double tmp = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      tmp += val[k] * x[colidx[k]];
    y[i] = 0.0 + tmp; /* dst.setZero(); dst += 1*tmp */
  }
}

/* Same per-row order, rows split over OpenMP threads: the multi-core CPU
 * baseline (each row is still summed sequentially, so results are identical). */
void or_spmv_f64_omp(int64_t nrows, const int32_t* rowptr,
                     const int32_t* colidx, const double* val, const double* x,
                     double* y, int nthreads)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
  for (int64_t i = 0; i < nrows; ++i)
  {
    double tmp = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      tmp += val[k] * x[colidx[k]];
    y[i] = 0.0 + tmp;
  }
  (void)nthreads;
}

/* Complex products (Matrix<std::complex<T>>, spmv/Matrix.cpp:441-444):
 * the same row order, std::complex's product without FMA
 * (re = ar xr - ai xi, im = ar xi + ai xr, each product and sum rounded).
 * Values and vectors are (re, im) interleaved. */
void or_spmv_c128(int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                  const double* val, const double* x, double* y)
{
  for (int64_t i = 0; i < nrows; ++i)
  {
    double tr = 0.0, ti = 0.0;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    {
      const double ar = val[2 * (int64_t)k], ai = val[2 * (int64_t)k + 1];
      const double xr = x[2 * (int64_t)colidx[k]], xi = x[2 * (int64_t)colidx[k] + 1];
      const double pr = ar * xr - ai * xi;
      const double pi = ar * xi + ai * xr;
      tr += pr;
      ti += pi;
    }
    y[2 * i] = tr;
    y[2 * i + 1] = ti;
  }
}

void or_spmv_c64(int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                 const float* val, const float* x, float* y)
{
  for (int64_t i = 0; i < nrows; ++i)
  {
    float tr = 0.0f, ti = 0.0f;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    {
      const float ar = val[2 * (int64_t)k], ai = val[2 * (int64_t)k + 1];
      const float xr = x[2 * (int64_t)colidx[k]], xi = x[2 * (int64_t)colidx[k] + 1];
      const float pr = ar * xr - ai * xi;
      const float pi = ar * xi + ai * xr;
      tr += pr;
      ti += pi;
    }
    y[2 * i] = tr;
    y[2 * i + 1] = ti;
  }
}

/* Complex transpose product (not conjugated): y[col_k] += a_k x_i, rows in
 * ascending order, as or_spmv_t_f64. */
void or_spmv_t_c128(int64_t nrows, int64_t ncols, const int32_t* rowptr,
                    const int32_t* colidx, const double* val, const double* x, double* y)
{
  for (int64_t j = 0; j < 2 * ncols; ++j)
    y[j] = 0.0;
  for (int64_t i = 0; i < nrows; ++i)
  {
    const double xr = x[2 * i], xi = x[2 * i + 1];
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    {
      const double ar = val[2 * (int64_t)k], ai = val[2 * (int64_t)k + 1];
      const double pr = ar * xr - ai * xi;
      const double pi = ar * xi + ai * xr;
      y[2 * (int64_t)colidx[k]] += pr;
      y[2 * (int64_t)colidx[k] + 1] += pi;
    }
  }
}

/* Transpose product y[ncols] = A^T x[nrows]: spmv/Matrix.cpp:206-211.
 * Eigen evaluates `_matA.transpose() * b` by scattering row i's entries:
 * y[col_k] += a_k * x_i, rows in ascending order. */
void or_spmv_t_f64(int64_t nrows, int64_t ncols, const int32_t* rowptr,
                   const int32_t* colidx, const double* val, const double* x,
                   double* y)
{
  for (int64_t j = 0; j < ncols; ++j)
    y[j] = 0.0;
  for (int64_t i = 0; i < nrows; ++i)
  {
    const double xi = x[i];
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      y[colidx[k]] += val[k] * xi;
  }
}

/* FP32 SpMV with FP32 accumulation (checker for the single-precision kernels). */
void or_spmv_f32(int64_t nrows, const int32_t* rowptr, const int32_t* colidx,
                 const float* val, const float* x, float* y)
{
  for (int64_t i = 0; i < nrows; ++i)
  {
    float tmp = 0.0f;
    for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
      tmp += val[k] * x[colidx[k]];
    y[i] = tmp;
  }
}

/* ------------------------------------------------------------------------ */
/* Ghost exchange on one rank (spmv/L2GMap.cpp:126-170).                     */
/* The neighbour transport is replaced by caller buffers:                    */
/*   update:   send[i] = v[indexbuf[i]]  (L2GMap.cpp:137-139)                */
/*             v[local_size + g] = halo[g] (what Neighbor_alltoallv lands    */
/*             in the contiguous ghost region, L2GMap.cpp:144-147)           */
/*   reverse:  out[g] = v[local_size + g] (sent to owners, :161-164)         */
/*             v[indexbuf[i]] += incoming[i]  (:168-169, in index order)     */
/* ------------------------------------------------------------------------ */
void or_l2g_update_f64(int64_t local_size, int64_t num_ghosts,
                       int64_t num_indices, const int32_t* indexbuf,
                       double* v, const double* halo, double* send)
{
  for (int64_t i = 0; i < num_indices; ++i)
    if (send)
      send[i] = v[indexbuf[i]];
  for (int64_t g = 0; g < num_ghosts; ++g)
    v[local_size + g] = halo ? halo[g] : 0.0;
}

void or_l2g_reverse_update_f64(int64_t local_size, int64_t num_ghosts,
                               int64_t num_indices, const int32_t* indexbuf,
                               double* v, const double* incoming,
                               double* ghost_out)
{
  for (int64_t g = 0; g < num_ghosts; ++g)
    if (ghost_out)
      ghost_out[g] = v[local_size + g];
  for (int64_t i = 0; i < num_indices; ++i)
    v[indexbuf[i]] += incoming[i];
}

/* ------------------------------------------------------------------------ */
/* CG: spmv/cg.cpp:11-83, line by line.                                      */
/*   ncols = local_size + num_ghosts of the column map; nrows = M.           */
/*   halo (length ncols-nrows, may be NULL = zeros) is what update(p)        */
/*   delivers into p's ghost region every iteration (cg.cpp:55).             */
/*   hist[0..k] receives rnorm_k (squared norms), hist[0] = rnorm0.          */
/* Returns k (iterations); x has length ncols with zero ghosts (cg.cpp:38).  */
/* ------------------------------------------------------------------------ */
static double or_dot(int64_t n, const double* a, const double* b)
{
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

/* Dot product in a fixed chunked order: the range is cut into OR_DOT_CHUNKS
 * contiguous chunks (the cut depends on n only), each chunk summed left to
 * right, then the chunk sums left to right.  Deterministic whatever the
 * number of threads; rounded differently from the sequential or_dot (the
 * reference's own order is Eigen's SIMD reduction, which neither restates). */
#define OR_DOT_CHUNKS 64
static double or_dot_chunked(int64_t n, const double* a, const double* b, int nthreads)
{
  double part[OR_DOT_CHUNKS];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
  for (int c = 0; c < OR_DOT_CHUNKS; ++c)
  {
    const int64_t lo = n * c / OR_DOT_CHUNKS, hi = n * (c + 1) / OR_DOT_CHUNKS;
    double s = 0.0;
    for (int64_t i = lo; i < hi; ++i)
      s += a[i] * b[i];
    part[c] = s;
  }
  (void)nthreads;
  double s = 0.0;
  for (int c = 0; c < OR_DOT_CHUNKS; ++c)
    s += part[c];
  return s;
}

/* The same loop with its product, vector updates and dot products split over
 * `nthreads` OpenMP threads, so that the checker can run CG on the
 * full-size systems (10^8 entries, 10^3 iterations) in tens of seconds.
 * Every row is still summed in stored order and every vector entry updated
 * by the same expression as in or_cg_f64; only the dot products use the
 * chunked order above.  The result does not depend on nthreads. */
int or_cg_f64_mt(int64_t nrows, int64_t ncols, const int32_t* rowptr,
                 const int32_t* colidx, const double* val, const double* b,
                 int kmax, double rtol, const double* halo, double* x,
                 double* hist, int nthreads)
{
  const int64_t M = nrows;
  if (nthreads < 1)
    nthreads = 1;
  double* r = (double*)malloc(sizeof(double) * (size_t)M);
  double* y = (double*)malloc(sizeof(double) * (size_t)M);
  double* p = (double*)calloc((size_t)ncols, sizeof(double));
  for (int64_t i = 0; i < ncols; ++i)
    x[i] = 0.0;                                  /* cg.cpp:39 */
  memcpy(r, b, sizeof(double) * (size_t)M);     /* cg.cpp:40 */
  memcpy(p, r, sizeof(double) * (size_t)M);     /* cg.cpp:41 */
  double rnorm0 = or_dot_chunked(M, r, r, nthreads); /* cg.cpp:42-44 */
  if (hist)
    hist[0] = rnorm0;
  const double rtol2 = rtol * rtol;
  double rnorm_old = rnorm0;
  int k = 0;
  while (k < kmax)
  {
    ++k;
    for (int64_t g = M; g < ncols; ++g)         /* update(p), cg.cpp:55 */
      p[g] = halo ? halo[g - M] : 0.0;
    or_spmv_f64_omp(M, rowptr, colidx, val, p, y, nthreads); /* cg.cpp:56 */
    double pdoty = or_dot_chunked(M, p, y, nthreads);        /* cg.cpp:59-61 */
    double alpha = rnorm_old / pdoty;           /* cg.cpp:62 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < M; ++i)
    {
      x[i] += alpha * p[i];                     /* cg.cpp:65 */
      r[i] -= alpha * y[i];                     /* cg.cpp:66 */
    }
    double rnorm_new = or_dot_chunked(M, r, r, nthreads); /* cg.cpp:69-71 */
    double beta = rnorm_new / rnorm_old;        /* cg.cpp:72 */
    rnorm_old = rnorm_new;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int64_t i = 0; i < M; ++i)
      p[i] = p[i] * beta + r[i];                /* cg.cpp:76 */
    if (hist)
      hist[k] = rnorm_new;
    if (rnorm_new / rnorm0 < rtol2)             /* cg.cpp:78-79 */
      break;
  }
  free(r);
  free(y);
  free(p);
  return k;
}

int or_cg_f64(int64_t nrows, int64_t ncols, const int32_t* rowptr,
              const int32_t* colidx, const double* val, const double* b,
              int kmax, double rtol, const double* halo, double* x,
              double* hist)
{
  const int64_t M = nrows;
  double* r = (double*)malloc(sizeof(double) * (size_t)M);
  double* y = (double*)malloc(sizeof(double) * (size_t)M);
  double* p = (double*)calloc((size_t)ncols, sizeof(double));
  for (int64_t i = 0; i < ncols; ++i)
    x[i] = 0.0;                                  /* cg.cpp:39 */
  memcpy(r, b, sizeof(double) * (size_t)M);     /* cg.cpp:40 */
  memcpy(p, r, sizeof(double) * (size_t)M);     /* cg.cpp:41 */
  double rnorm0 = or_dot(M, r, r);              /* cg.cpp:42-44 */
  if (hist)
    hist[0] = rnorm0;
  const double rtol2 = rtol * rtol;
  double rnorm_old = rnorm0;
  int k = 0;
  while (k < kmax)
  {
    ++k;
    for (int64_t g = M; g < ncols; ++g)         /* update(p), cg.cpp:55 */
      p[g] = halo ? halo[g - M] : 0.0;
    or_spmv_f64(M, rowptr, colidx, val, p, y);  /* cg.cpp:56 */
    double pdoty = or_dot(M, p, y);             /* cg.cpp:59-61 */
    double alpha = rnorm_old / pdoty;           /* cg.cpp:62 */
    for (int64_t i = 0; i < M; ++i)
      x[i] += alpha * p[i];                     /* cg.cpp:65 */
    for (int64_t i = 0; i < M; ++i)
      r[i] -= alpha * y[i];                     /* cg.cpp:66 */
    double rnorm_new = or_dot(M, r, r);         /* cg.cpp:69-71 */
    double beta = rnorm_new / rnorm_old;        /* cg.cpp:72 */
    rnorm_old = rnorm_new;
    for (int64_t i = 0; i < M; ++i)
      p[i] = p[i] * beta + r[i];                /* cg.cpp:76 */
    if (hist)
      hist[k] = rnorm_new;
    if (rnorm_new / rnorm0 < rtol2)             /* cg.cpp:78-79 */
      break;
  }
  free(r);
  free(y);
  free(p);
  return k;
}

/* FP32 CG (same loop, single precision throughout): checker for the FP32
 * inner solver and the batched solver.  max_restarts > 0 adds the batched
 * solver's true-residual check: when the recursive residual meets rtol, the
 * true residual b - A x is formed with products and sums in double (the FP32
 * values and x are exact in double); if it does not meet rtol either, CG
 * restarts from x with r = p = (float)(b - A x) and ||r||^2 = (float) of its
 * double norm, and a stopping threshold four times lower, at most
 * max_restarts times.  Iterations count on across restarts. */
int or_cg_f32_restart(int64_t n, const int32_t* rowptr, const int32_t* colidx,
                      const float* val, const float* b, int kmax, float rtol,
                      float* x, float* hist, int max_restarts)
{
  float* r = (float*)malloc(sizeof(float) * (size_t)n);
  float* y = (float*)malloc(sizeof(float) * (size_t)n);
  float* p = (float*)malloc(sizeof(float) * (size_t)n);
  for (int64_t i = 0; i < n; ++i)
  {
    x[i] = 0.0f;
    r[i] = b[i];
    p[i] = b[i];
  }
  float rnorm0 = 0.0f;
  double bnorm2 = 0.0;
  for (int64_t i = 0; i < n; ++i)
  {
    rnorm0 += r[i] * r[i];
    bnorm2 += (double)b[i] * (double)b[i];
  }
  if (hist)
    hist[0] = rnorm0;
  const float rtol2 = rtol * rtol;
  float tgt2 = rtol2; /* a quarter of it again after every failed check */
  float rnorm_old = rnorm0;
  int k = 0, restarts = 0;
  while (k < kmax)
  {
    ++k;
    or_spmv_f32(n, rowptr, colidx, val, p, y);
    float pdoty = 0.0f;
    for (int64_t i = 0; i < n; ++i)
      pdoty += p[i] * y[i];
    float alpha = rnorm_old / pdoty;
    float rnorm_new = 0.0f;
    for (int64_t i = 0; i < n; ++i)
    {
      x[i] += alpha * p[i];
      r[i] -= alpha * y[i];
      rnorm_new += r[i] * r[i];
    }
    float beta = rnorm_new / rnorm_old;
    rnorm_old = rnorm_new;
    for (int64_t i = 0; i < n; ++i)
      p[i] = p[i] * beta + r[i];
    if (hist)
      hist[k] = rnorm_new;
    if (rnorm_new / rnorm0 < tgt2)
    {
      if (restarts >= max_restarts)
        break;
      double rt2 = 0.0;
      for (int64_t i = 0; i < n; ++i)
      {
        double t = 0.0;
        for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
          t += (double)val[e] * (double)x[colidx[e]];
        const double ri = (double)b[i] - t;
        rt2 += ri * ri;
        r[i] = (float)ri;
        p[i] = r[i];
      }
      if (rt2 < (double)rtol * (double)rtol * bnorm2)
        break;
      ++restarts;
      rnorm_old = (float)rt2;
      tgt2 *= 0.25f;
    }
  }
  free(r);
  free(y);
  free(p);
  return k;
}

int or_cg_f32(int64_t n, const int32_t* rowptr, const int32_t* colidx,
              const float* val, const float* b, int kmax, float rtol,
              float* x, float* hist)
{
  return or_cg_f32_restart(n, rowptr, colidx, val, b, kmax, rtol, x, hist, 0);
}

/* ------------------------------------------------------------------------ */
/* Generators.  Grid point (i,j,k) -> row i + nx*(j + ny*k).  Columns are    */
/* emitted in ascending order per row (what setFromTriplets/makeCompressed   */
/* produce, Matrix.cpp:427-429, CreateA.cpp:75).  Each returns nnz; call     */
/* with NULL arrays first to size them.                                      */
/* ------------------------------------------------------------------------ */

/* 1D demo matrix of demos/CreateA.cpp:52-74 on one rank (global = local). */
int64_t or_gen_demo1d(int64_t N, double gamma, int32_t* rowptr,
                      int32_t* colidx, double* val)
{
  int64_t nnz = 0;
  if (rowptr)
    rowptr[0] = 0;
  for (int64_t c0 = 0; c0 < N; ++c0)
  {
    int64_t cols[3];
    double vals[3];
    int n = 0;
    if (c0 == 0)
    {
      cols[n] = c0;     vals[n++] = 1.0 - gamma;
      if (N > 1) { cols[n] = c0 + 1; vals[n++] = gamma; }
    }
    else if (c0 == N - 1)
    {
      cols[n] = c0 - 1; vals[n++] = gamma;
      cols[n] = c0;     vals[n++] = 1.0 - gamma;
    }
    else
    {
      cols[n] = c0 - 1; vals[n++] = gamma;
      cols[n] = c0;     vals[n++] = 1.0 - 2.0 * gamma;
      cols[n] = c0 + 1; vals[n++] = gamma;
    }
    for (int t = 0; t < n; ++t, ++nnz)
      if (colidx)
      {
        colidx[nnz] = (int32_t)cols[t];
        val[nnz] = vals[t];
      }
    if (rowptr)
      rowptr[c0 + 1] = (int32_t)nnz;
  }
  return nnz;
}

/* Generic box stencil on an nx*ny*nz grid with Dirichlet boundaries.
 *   kind 5: 2D 5-point (nz must be 1), 7: 3D 7-point, 27: 3D 27-point.
 *   diag on the diagonal, off on every in-grid neighbour.
 *   ghost_planes: number of extra z-planes above the top (k = nz .. nz+gp-1)
 *   that exist as ghost columns: numbered ncols_owned + (i + nx*(j + ny*g)),
 *   i.e. ascending global order of the plane owned by the next rank.
 */
int64_t or_gen_stencil(int kind, int64_t nx, int64_t ny, int64_t nz,
                       int64_t ghost_planes, double diag, double off,
                       int32_t* rowptr, int32_t* colidx, double* val)
{
  const int64_t nowned = nx * ny * nz;
  int64_t nnz = 0;
  if (rowptr)
    rowptr[0] = 0;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
      {
        const int64_t row = i + nx * (j + ny * k);
        for (int dk = -1; dk <= 1; ++dk)
          for (int dj = -1; dj <= 1; ++dj)
            for (int di = -1; di <= 1; ++di)
            {
              const int nzero = (di != 0) + (dj != 0) + (dk != 0);
              if (kind == 5 && (dk != 0 || nzero > 1))
                continue;
              if (kind == 7 && nzero > 1)
                continue;
              const int64_t ii = i + di, jj = j + dj, kk = k + dk;
              if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0)
                continue;
              if (kk >= nz + ghost_planes)
                continue;
              const int64_t col = ii + nx * (jj + ny * kk);
              if (colidx)
              {
                colidx[nnz] = (int32_t)col; /* ghosts land at nowned + ... */
                val[nnz] = (nzero == 0) ? diag : off;
              }
              ++nnz;
            }
        if (rowptr)
          rowptr[row + 1] = (int32_t)nnz;
      }
  (void)nowned;
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* 2D anisotropic diffusion on an n*n grid, edge conductivities theta.       */
/* Edges: horizontal (i,j)-(i+1,j): index j*(n-1)+i, count n*(n-1);          */
/*        vertical   (i,j)-(i,j+1): index n*(n-1) + j*n + i, count n*(n-1);  */
/*        boundary   one per boundary side of a boundary node, 4n in total:  */
/*        index 2n(n-1) + side*n + t (side 0:j=0 row, 1:j=n-1, 2:i=0, 3:i=n-1)*/
/* A = sum_interior theta_e (e_a-e_b)(e_a-e_b)^T + sum_bnd theta_e e_a e_a^T  */
/* (Dirichlet: a boundary edge connects node a to a zero-valued outside).    */
/* ------------------------------------------------------------------------ */
int64_t or_diffusion_num_edges(int64_t n) { return 2 * n * (n - 1) + 4 * n; }

int64_t or_gen_diffusion(int64_t n, const double* theta, int32_t* rowptr,
                         int32_t* colidx, double* val)
{
  const int64_t nh = n * (n - 1);
  int64_t nnz = 0;
  if (rowptr)
    rowptr[0] = 0;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < n; ++i)
    {
      const int64_t row = i + n * j;
      /* edges touching this node */
      const int64_t e_s = (j > 0) ? nh + (j - 1) * n + i : -1;     /* below */
      const int64_t e_w = (i > 0) ? j * (n - 1) + (i - 1) : -1;    /* left  */
      const int64_t e_e = (i < n - 1) ? j * (n - 1) + i : -1;      /* right */
      const int64_t e_n = (j < n - 1) ? nh + j * n + i : -1;       /* above */
      double d = 0.0;
      if (theta)
      {
        /* ascending edge index: w < e < s < n < boundary edges */
        if (e_w >= 0) d += theta[e_w];
        if (e_e >= 0) d += theta[e_e];
        if (e_s >= 0) d += theta[e_s];
        if (e_n >= 0) d += theta[e_n];
        if (j == 0) d += theta[2 * nh + 0 * n + i];
        if (j == n - 1) d += theta[2 * nh + 1 * n + i];
        if (i == 0) d += theta[2 * nh + 2 * n + j];
        if (i == n - 1) d += theta[2 * nh + 3 * n + j];
      }
      const int64_t cols[5] = {row - n, row - 1, row, row + 1, row + n};
      const int64_t edges[5] = {e_s, e_w, -2, e_e, e_n};
      for (int t = 0; t < 5; ++t)
      {
        if (edges[t] == -1)
          continue;
        if (colidx)
        {
          colidx[nnz] = (int32_t)cols[t];
          val[nnz] = (edges[t] == -2) ? d : (theta ? -theta[edges[t]] : 0.0);
        }
        ++nnz;
      }
      if (rowptr)
        rowptr[row + 1] = (int32_t)nnz;
    }
  return nnz;
}

/* Gradient of a scalar loss wrt theta from grad_A on the pattern:
 *   grad_theta_e = G_aa + G_bb - G_ab - G_ba  (interior edge a-b)
 *   grad_theta_e = G_aa                       (boundary edge at a)
 * with G given as values on the CSR pattern of or_gen_diffusion. */
void or_diffusion_grad_theta(int64_t n, const int32_t* rowptr,
                             const int32_t* colidx, const double* G,
                             double* grad_theta)
{
  const int64_t nh = n * (n - 1);
  const int64_t ne = or_diffusion_num_edges(n);
  for (int64_t e = 0; e < ne; ++e)
    grad_theta[e] = 0.0;
  for (int64_t row = 0; row < n * n; ++row)
  {
    const int64_t i = row % n, j = row / n;
    double gdiag = 0.0;
    for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k)
      if (colidx[k] == row)
        gdiag = G[k];
    /* every edge touching row gets +G_aa */
    if (j > 0) grad_theta[nh + (j - 1) * n + i] += gdiag;
    if (i > 0) grad_theta[j * (n - 1) + (i - 1)] += gdiag;
    if (i < n - 1) grad_theta[j * (n - 1) + i] += gdiag;
    if (j < n - 1) grad_theta[nh + j * n + i] += gdiag;
    if (j == 0) grad_theta[2 * nh + 0 * n + i] += gdiag;
    if (j == n - 1) grad_theta[2 * nh + 1 * n + i] += gdiag;
    if (i == 0) grad_theta[2 * nh + 2 * n + j] += gdiag;
    if (i == n - 1) grad_theta[2 * nh + 3 * n + j] += gdiag;
    /* off-diagonal entries: -G_ab for the edge a-b */
    for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k)
    {
      const int64_t c = colidx[k];
      int64_t e = -1;
      if (c == row - n) e = nh + (j - 1) * n + i;
      else if (c == row - 1) e = j * (n - 1) + (i - 1);
      else if (c == row + 1) e = j * (n - 1) + i;
      else if (c == row + n) e = nh + j * n + i;
      if (e >= 0)
        grad_theta[e] -= G[k];
    }
  }
}

/* Adjoint of L = 0.5||x - x_t||^2 for A(theta) x = b (A symmetric, so the
 * adjoint system A^T lam = x - x_t is solved with the same CG):
 *   grad_b = lam, grad_A = -lam x^T on the pattern, grad_theta by the edge
 *   chain rule above.  Returns total CG iterations (forward + adjoint). */
int or_adjoint_diffusion(int64_t n, const double* theta, const double* b,
                         const double* x_target, int kmax, double rtol,
                         double* x, double* lam, double* grad_theta,
                         double* loss)
{
  const int64_t N = n * n;
  int64_t nnz = or_gen_diffusion(n, NULL, NULL, NULL, NULL);
  int32_t* rowptr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(N + 1));
  int32_t* colidx = (int32_t*)malloc(sizeof(int32_t) * (size_t)nnz);
  double* val = (double*)malloc(sizeof(double) * (size_t)nnz);
  double* G = (double*)malloc(sizeof(double) * (size_t)nnz);
  double* rhs = (double*)malloc(sizeof(double) * (size_t)N);
  or_gen_diffusion(n, theta, rowptr, colidx, val);
  int k1 = or_cg_f64(N, N, rowptr, colidx, val, b, kmax, rtol, NULL, x, NULL);
  double L = 0.0;
  for (int64_t i = 0; i < N; ++i)
  {
    rhs[i] = x[i] - x_target[i];
    L += 0.5 * rhs[i] * rhs[i];
  }
  if (loss)
    *loss = L;
  int k2 = or_cg_f64(N, N, rowptr, colidx, val, rhs, kmax, rtol, NULL, lam,
                     NULL);
  for (int64_t row = 0; row < N; ++row)
    for (int32_t k = rowptr[row]; k < rowptr[row + 1]; ++k)
      G[k] = -lam[row] * x[colidx[k]];
  or_diffusion_grad_theta(n, rowptr, colidx, G, grad_theta);
  free(rowptr);
  free(colidx);
  free(val);
  free(G);
  free(rhs);
  return k1 + k2;
}
