// Device kernels of libspmv_amd (gfx950).
//
// CSR SpMV, "ordered stream" kernel
// ---------------------------------
// The analysis (matrix.hip) cuts the rows into blocks of at most kBS rows
// holding at most kChunk stored entries (or one longer row).  One workgroup
// per block:
//   phase 1  all kBS threads walk the block's contiguous entry range with
//            coalesced loads of colidx/values, gather x[colidx], and store the
//            rounded products a_k * x_c into LDS;
//   phase 2  thread t sums row r0+t's products from LDS sequentially in
//            stored order, starting from 0.
// Each product and each sum is rounded separately (no contraction is
// possible: the operands of the sums come from LDS), which is exactly the
// Eigen 3.3 row-major product of the reference (spmv/Matrix.cpp:203): the
// FP64 result is bit-identical to it, whatever the row lengths.
// Workgroups are mapped XCD-contiguously (xcd_block) so the rows processed by
// one XCD are adjacent and share x entries in that XCD's L2.
#pragma once

#include "common.hpp"

namespace spmvx
{

constexpr int kBS = 256;     // threads per SpMV workgroup = max rows per block
constexpr int kChunk = 2048; // products staged in LDS per block

// ---------------------------------------------------------------- values
struct bf16_t
{
  uint16_t bits;
};

__device__ __forceinline__ double load_val(const double* v, int64_t k) { return v[k]; }
__device__ __forceinline__ float load_val(const float* v, int64_t k) { return v[k]; }
__device__ __forceinline__ float load_val(const bf16_t* v, int64_t k)
{
  return __uint_as_float(static_cast<uint32_t>(v[k].bits) << 16);
}

// ---------------------------------------------------------------- CG state
// Scalars of one CG solve, resident on the device.  `active` is read at the
// start of every kernel and only changed by the last block of the final
// kernel of an iteration, so all blocks of a kernel agree on it.
template <typename Acc>
struct CgState
{
  Acc rnorm0;
  Acc rnorm_old;
  Acc alpha;
  Acc beta;
  Acc rtol2;
  Acc last;    // last reduction result (residual kernels)
  int32_t k;
  int32_t kmax;
  int32_t active;
  int32_t converged;
};


// Deterministic two-level grid reduction.  Every block calls it once with its
// block-level sum `v` (valid in thread 0) and its logical index `b` of `nb`.
// Level 1: the last block to arrive in each group of 256 blocks sums that
// group's partials; level 2: the last group finisher sums the group results.
// Returns true (in all threads of that one block) with the total in *total.
// partials: nb + ceil(nb/256) entries; counters: 1 + ceil(nb/256) entries,
// zero on first use (each counter is reset by its last arriver).
template <typename Acc>
__device__ __forceinline__ bool grid_reduce(Acc v, int64_t b, int64_t nb, Acc* partials,
                                            unsigned* counters, Acc* red, Acc* total)
{
  constexpr int64_t G = 256;
  const int tid = threadIdx.x;
  const int64_t ng = (nb + G - 1) / G;
  const int64_t g = b / G;
  if (tid == 0)
    partials[b] = v;
  const int64_t gsize = (nb - g * G) < G ? (nb - g * G) : G;
  if (!last_block_arrive(&counters[1 + g], static_cast<unsigned>(gsize)))
    return false;
  Acc s = Acc(0);
  for (int64_t i = tid; i < gsize; i += blockDim.x)
    s += partials[g * G + i];
  s = block_sum(s, red);
  if (tid == 0)
    partials[nb + g] = s;
  if (!last_block_arrive(&counters[0], static_cast<unsigned>(ng)))
    return false;
  Acc t = Acc(0);
  for (int64_t i = tid; i < ng; i += blockDim.x)
    t += partials[nb + i];
  t = block_sum(t, red);
  *total = t;
  return true;
}

enum SpmvMode
{
  kPlain = 0, // y = A x
  kDot = 1,   // y = A p, alpha = rnorm_old / (p . y)     (CG, cg.cpp:56-62)
  kResid = 2  // y = b - A x, last = ||y||^2              (refinement)
};

// Reduction workspace of one kernel: partials[nb + nb/256 + 1], counters[...].
template <typename Acc>
struct RedWs
{
  Acc* partials;
  unsigned* counters;
};

template <typename V, typename X, typename Acc, int MODE>
__global__ __launch_bounds__(kBS) void csr_stream(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const V* __restrict__ val, const X* __restrict__ x, X* __restrict__ y,
    const int32_t* __restrict__ blocks, int64_t nblocks, const X* __restrict__ b,
    RedWs<Acc> ws, CgState<Acc>* st)
{
  if (MODE != kPlain && !st->active)
    return;
  __shared__ Acc prod[kChunk];
  __shared__ Acc red[16];

  const int64_t bi = xcd_block(blockIdx.x, nblocks);
  const int32_t r0 = blocks[bi], r1 = blocks[bi + 1];
  const int32_t k0 = rowptr[r0], k1 = rowptr[r1];
  const int tid = threadIdx.x;
  Acc part = Acc(0);

  if (k1 - k0 <= kChunk)
  {
    const int n = k1 - k0;
    for (int t = tid; t < n; t += kBS)
    {
      const int64_t k = k0 + t;
      prod[t] = static_cast<Acc>(load_val(val, k)) * static_cast<Acc>(x[colidx[k]]);
    }
    __syncthreads();
    const int32_t row = r0 + tid;
    if (row < r1)
    {
      const int a = rowptr[row] - k0, e = rowptr[row + 1] - k0;
      Acc s = Acc(0);
      for (int q = a; q < e; ++q)
        s += prod[q];
      if (MODE == kResid)
      {
        const X yi = b[row] - static_cast<X>(s);
        y[row] = yi;
        part = static_cast<Acc>(yi) * static_cast<Acc>(yi);
      }
      else
      {
        const X yi = static_cast<X>(s);
        y[row] = yi;
        if (MODE == kDot)
          part = static_cast<Acc>(x[row]) * static_cast<Acc>(yi);
      }
    }
  }
  else
  {
    // One row longer than kChunk: staged chunk by chunk, summed in order by
    // thread 0.
    Acc s = Acc(0);
    for (int64_t c0 = k0; c0 < k1; c0 += kChunk)
    {
      const int n = static_cast<int>((k1 - c0) < kChunk ? (k1 - c0) : kChunk);
      __syncthreads();
      for (int t = tid; t < n; t += kBS)
      {
        const int64_t k = c0 + t;
        prod[t] = static_cast<Acc>(load_val(val, k)) * static_cast<Acc>(x[colidx[k]]);
      }
      __syncthreads();
      if (tid == 0)
        for (int q = 0; q < n; ++q)
          s += prod[q];
    }
    if (tid == 0)
    {
      if (MODE == kResid)
      {
        const X yi = b[r0] - static_cast<X>(s);
        y[r0] = yi;
        part = static_cast<Acc>(yi) * static_cast<Acc>(yi);
      }
      else
      {
        const X yi = static_cast<X>(s);
        y[r0] = yi;
        if (MODE == kDot)
          part = static_cast<Acc>(x[r0]) * static_cast<Acc>(yi);
      }
    }
  }

  if (MODE == kPlain)
    return;
  part = block_sum(part, red);
  Acc total;
  if (grid_reduce(part, bi, nblocks, ws.partials, ws.counters, red, &total) && tid == 0)
  {
    if (MODE == kDot)
      st->alpha = st->rnorm_old / total;
    else
      st->last = total;
  }
}

// ---------------------------------------------------------------- CG vectors
constexpr int kVecBS = 256;

// x = 0, r = b, p[0:n) = b, p[n:ncols) = 0, rnorm0 = ||b||^2 (cg.cpp:31-44)
template <typename X, typename Acc>
__global__ __launch_bounds__(kVecBS) void cg_init(int64_t n, int64_t ncols,
                                                  const X* __restrict__ b, X* __restrict__ x,
                                                  X* __restrict__ r, X* __restrict__ p,
                                                  RedWs<Acc> ws, CgState<Acc>* st)
{
  __shared__ Acc red[16];
  Acc part = Acc(0);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBS;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kVecBS) + threadIdx.x; i < ncols; i += stride)
  {
    x[i] = X(0);
    if (i < n)
    {
      const X bi = b[i];
      r[i] = bi;
      p[i] = bi;
      part += static_cast<Acc>(bi) * static_cast<Acc>(bi);
    }
    else
      p[i] = X(0);
  }
  part = block_sum(part, red);
  Acc total;
  if (grid_reduce(part, blockIdx.x, gridDim.x, ws.partials, ws.counters, red, &total)
      && threadIdx.x == 0)
  {
    st->rnorm0 = total;
    st->rnorm_old = total;
    st->k = 0;
    st->converged = 0;
    st->active = st->kmax > 0 ? 1 : 0;
  }
}

// r -= alpha y; rnorm_new = ||r||^2; beta; k++; convergence (cg.cpp:66-79)
template <typename X, typename Acc>
__global__ __launch_bounds__(kVecBS) void cg_update_r(int64_t n, X* __restrict__ r,
                                                      const X* __restrict__ y, RedWs<Acc> ws,
                                                      CgState<Acc>* st, Acc* __restrict__ hist)
{
  if (!st->active)
    return;
  __shared__ Acc red[16];
  const X alpha = static_cast<X>(st->alpha);
  Acc part = Acc(0);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBS;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kVecBS) + threadIdx.x; i < n; i += stride)
  {
    const X ri = r[i] - alpha * y[i];
    r[i] = ri;
    part += static_cast<Acc>(ri) * static_cast<Acc>(ri);
  }
  part = block_sum(part, red);
  Acc s;
  if (grid_reduce(part, blockIdx.x, gridDim.x, ws.partials, ws.counters, red, &s)
      && threadIdx.x == 0)
  {
    st->beta = s / st->rnorm_old;
    st->rnorm_old = s;
    st->k += 1;
    if (hist)
      hist[st->k] = s;
    st->converged = (s / st->rnorm0 < st->rtol2) ? 1 : 0;
  }
}

// x += alpha p; p = beta p + r; stop after this iteration if converged or
// k == kmax (cg.cpp:65,76-79)
template <typename X, typename Acc>
__global__ __launch_bounds__(kVecBS) void cg_update_xp(int64_t n, X* __restrict__ x,
                                                       X* __restrict__ p,
                                                       const X* __restrict__ r,
                                                       unsigned* counter, CgState<Acc>* st)
{
  if (!st->active)
    return;
  const X alpha = static_cast<X>(st->alpha);
  const X beta = static_cast<X>(st->beta);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBS;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kVecBS) + threadIdx.x; i < n; i += stride)
  {
    const X pi = p[i];
    x[i] = x[i] + alpha * pi;
    p[i] = pi * beta + r[i];
  }
  if (last_block_arrive(counter, gridDim.x) && threadIdx.x == 0)
  {
    if (st->converged || st->k >= st->kmax)
      st->active = 0;
  }
}

} // namespace spmvx
