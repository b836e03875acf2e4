// Batched CG: many independent SPD systems (the diagonal blocks of one
// block-diagonal CSR matrix) solved by ONE persistent kernel launch, captured
// in a hipGraph.  No reference counterpart; each system runs the loop of
// spmv::cg (spmv/cg.cpp:50-80) in FP32 with NRHS right-hand sides at once.
//
// Design (gfx950):
//   * one workgroup solves one whole system (all right-hand sides); all of
//     its synchronisation is workgroup-local (__syncthreads), so workgroups
//     never wait on each other and the grid always drains;
//   * systems are pulled from a device work queue (atomic counter) in
//     decreasing size order, so the largest start first (LPT schedule);
//   * vectors are stored row-major [row][NRHS]: one 16-byte gather of p[c]
//     serves all four right-hand sides, and the matrix is read once per
//     iteration for all of them;
//   * G lanes cooperate on a row (rows per wavefront = 64/G), chosen per
//     system; G and the block size are picked by the autotuner.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>

#include "handles.hpp"

namespace spmvx
{
namespace
{

constexpr int kMaxBatchBS = 1024;

template <int NRHS>
struct VecN
{
  float v[NRHS];
};

template <int NRHS>
__device__ __forceinline__ VecN<NRHS> ldn(const float* p)
{
  VecN<NRHS> r;
  if constexpr (NRHS == 4)
  {
    const float4 q = *reinterpret_cast<const float4*>(p);
    r.v[0] = q.x;
    r.v[1] = q.y;
    r.v[2] = q.z;
    r.v[3] = q.w;
  }
  else if constexpr (NRHS == 2)
  {
    const float2 q = *reinterpret_cast<const float2*>(p);
    r.v[0] = q.x;
    r.v[1] = q.y;
  }
  else
    r.v[0] = *p;
  return r;
}

template <int NRHS>
__device__ __forceinline__ void stn(float* p, const VecN<NRHS>& r)
{
  if constexpr (NRHS == 4)
    *reinterpret_cast<float4*>(p) = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  else if constexpr (NRHS == 2)
    *reinterpret_cast<float2*>(p) = make_float2(r.v[0], r.v[1]);
  else
    *p = r.v[0];
}

// Sum over the block of NRHS values per thread; result in all threads.
template <int NRHS>
__device__ __forceinline__ VecN<NRHS> block_sum_n(VecN<NRHS> a, float (*red)[16])
{
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int j = 0; j < NRHS; ++j)
    a.v[j] = wave_sum(a.v[j]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      red[j][wid] = a.v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NRHS; ++j)
    a.v[j] = wave_sum(lane < nw ? red[j][lane] : 0.0f);
  return a;
}

template <int NRHS>
struct VecD
{
  double v[NRHS];
};

// Sum over the block of NRHS doubles per thread; result in all threads.
template <int NRHS>
__device__ __forceinline__ VecD<NRHS> block_sum_d(VecD<NRHS> a, double (*red)[16])
{
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int j = 0; j < NRHS; ++j)
    a.v[j] = wave_sum(a.v[j]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      red[j][wid] = a.v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NRHS; ++j)
    a.v[j] = wave_sum(lane < nw ? red[j][lane] : 0.0);
  return a;
}

// True residual over rows [r0, r1), G lanes per row: t = b - A x with the
// caller's b read through perm (when the batch is reordered), the products
// and sums in double (the FP32 values and x are exact in double),
// r = p = (float) t; returns this thread's share of ||t||^2 (group leaders).
template <int NRHS, int G>
__device__ __forceinline__ VecD<NRHS> batch_resid(const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ colidx,
                                                  const float* __restrict__ val,
                                                  const int32_t* __restrict__ perm,
                                                  const float* __restrict__ b,
                                                  const float* __restrict__ x,
                                                  float* __restrict__ r, float* __restrict__ p,
                                                  int32_t r0, int32_t r1)
{
  VecD<NRHS> tt{};
  const int gl = threadIdx.x % G;
  const int rows_per_pass = blockDim.x / G;
  for (int32_t row = r0 + static_cast<int32_t>(threadIdx.x / G); row < r1; row += rows_per_pass)
  {
    VecD<NRHS> acc{};
    const int32_t e0 = rowptr[row], e1 = rowptr[row + 1];
    for (int32_t e = e0 + gl; e < e1; e += G)
    {
      const double a = static_cast<double>(val[e]);
      const VecN<NRHS> xc = ldn<NRHS>(x + static_cast<int64_t>(colidx[e]) * NRHS);
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
        acc.v[j] += a * static_cast<double>(xc.v[j]);
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
        acc.v[j] += __shfl_xor(acc.v[j], off, 64);
    if (gl == 0)
    {
      const int64_t o = static_cast<int64_t>(row) * NRHS;
      const VecN<NRHS> bi = ldn<NRHS>(b + static_cast<int64_t>(perm ? perm[row] : row) * NRHS);
      VecN<NRHS> ri;
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
      {
        const double t = static_cast<double>(bi.v[j]) - acc.v[j];
        tt.v[j] += t * t;
        ri.v[j] = static_cast<float>(t);
      }
      stn<NRHS>(r + o, ri);
      stn<NRHS>(p + o, ri);
    }
  }
  return tt;
}

// y = A p over rows [r0, r1) with G lanes per row; returns this thread's
// share of p.y (non-zero only in group leaders).  Each lane loads up to U
// (colidx, value) pairs, then issues their U gathers of p, before
// accumulating: U independent memory requests in flight per lane.
template <int NRHS, int G>
__device__ __forceinline__ VecN<NRHS> batch_spmv(const int32_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ colidx,
                                                 const float* __restrict__ val,
                                                 const float* __restrict__ p,
                                                 float* __restrict__ y, int32_t r0, int32_t r1)
{
  constexpr int U = 4;
  VecN<NRHS> pdy{};
  const int gl = threadIdx.x % G;
  const int rows_per_pass = blockDim.x / G;
  for (int32_t row = r0 + static_cast<int32_t>(threadIdx.x / G); row < r1; row += rows_per_pass)
  {
    VecN<NRHS> acc{};
    const int32_t e0 = rowptr[row], e1 = rowptr[row + 1];
    for (int32_t e = e0 + gl; e < e1; e += U * G)
    {
      int32_t c[U];
      float a[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
        const int32_t k = e + u * G;
        c[u] = k < e1 ? colidx[k] : row;
        a[u] = k < e1 ? val[k] : 0.0f;
      }
      VecN<NRHS> pc[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
        pc[u] = ldn<NRHS>(p + static_cast<int64_t>(c[u]) * NRHS);
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int j = 0; j < NRHS; ++j)
          acc.v[j] += a[u] * pc[u].v[j];
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
        acc.v[j] += __shfl_xor(acc.v[j], off, 64);
    if (gl == 0)
    {
      stn<NRHS>(y + static_cast<int64_t>(row) * NRHS, acc);
      const VecN<NRHS> pr = ldn<NRHS>(p + static_cast<int64_t>(row) * NRHS);
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
        pdy.v[j] += pr.v[j] * acc.v[j];
    }
  }
  return pdy;
}

// Bounded spin: a team barrier gives up (and flags the solve as failed)
// after about this many polls (a few seconds) instead of waiting forever; a
// legitimate wait lasts at most one team member's solve of another system.
constexpr unsigned kSpinLimit = 1u << 22;

// Barrier among the W workgroups of one team (all of them are running: a
// workgroup joins a team only by taking one of its slots from the queue).
// Arrivals are counted on a per-system counter; the t-th barrier of a solve
// completes at W * t arrivals.  An agent-scope release before the arrival
// and an acquire after the wait make the vector slices and partial sums
// written before the barrier visible to every member (members may run on
// different XCDs, whose L2s are not coherent with each other: the release
// writes this XCD's L2 back, the acquire invalidates it; a full fence on
// both sides would do each twice).
__device__ __forceinline__ void team_barrier(unsigned* bar, unsigned target, int* err)
{
  __syncthreads();
  if (threadIdx.x == 0)
  {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    atomicAdd(bar, 1u);
    unsigned spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
    {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kSpinLimit)
      {
        atomicExch(err, 1);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Team-wide sum: member m stores its block sum in slot m of buffer `which`
// (two buffers alternate), then every member adds the W slots in member
// order (identical result in all members).
template <int NRHS>
__device__ __forceinline__ VecN<NRHS> team_sum(VecN<NRHS> v, float* tred, int which, int W, int m,
                                               unsigned* bar, unsigned& nbar, int* err,
                                               float (*red)[16])
{
  v = block_sum_n<NRHS>(v, red);
  if (W == 1)
    return v;
  float* buf = tred + static_cast<int64_t>(which) * W * NRHS;
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      buf[m * NRHS + j] = v.v[j];
  team_barrier(bar, static_cast<unsigned>(W) * (++nbar), err);
  VecN<NRHS> t{};
  for (int q = 0; q < W; ++q)
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      t.v[j] += buf[q * NRHS + j];
  return t;
}

// team_sum in double (buffers of W * NRHS doubles, two alternating).
template <int NRHS>
__device__ __forceinline__ VecD<NRHS> team_sum_d(VecD<NRHS> v, double* tredd, int which, int W,
                                                 int m, unsigned* bar, unsigned& nbar, int* err,
                                                 double (*red)[16])
{
  v = block_sum_d<NRHS>(v, red);
  if (W == 1)
    return v;
  double* buf = tredd + static_cast<int64_t>(which) * W * NRHS;
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      buf[m * NRHS + j] = v.v[j];
  team_barrier(bar, static_cast<unsigned>(W) * (++nbar), err);
  VecD<NRHS> t{};
  for (int q = 0; q < W; ++q)
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      t.v[j] += buf[q * NRHS + j];
  return t;
}

// Index of the XCD (accelerator complex die) the calling wave runs on, from
// the XCC_ID hardware register.  Used only to pick a work queue: a wrong
// value costs locality, never correctness.
__device__ __forceinline__ int current_xcd()
{
  // s_getreg(HW_REG_XCC_ID = 20, offset 0, 4 bits)
  return static_cast<int>(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20)) & 0xF;
}

// True-residual checks per solve (each costs one product in double).
constexpr int kMaxRestarts = 4;

template <int NRHS>
__global__ __launch_bounds__(kMaxBatchBS) void batch_cg_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const float* __restrict__ val, const int32_t* __restrict__ sys_row,
    const int2* __restrict__ slots, const int32_t* __restrict__ qoff, int32_t nq,
    const int32_t* __restrict__ team, const int64_t* __restrict__ tred_off,
    const int8_t* __restrict__ gsel, const int32_t* __restrict__ perm,
    const float* __restrict__ b, float* __restrict__ xu, float* __restrict__ xin,
    float* __restrict__ r, float* __restrict__ p, float* __restrict__ y, int32_t max_its,
    float rtol2, double rtol2d, int32_t* __restrict__ its_out, unsigned* queue, unsigned* bars,
    float* tred_all, int* err)
{
  __shared__ int2 job;
  __shared__ float red[NRHS][16];
  __shared__ double redd[NRHS][16];
  // queues are per XCD: members of one team share an L2; a workgroup whose
  // own queue is empty takes slots from the next ones
  const int home = current_xcd() % nq;
  int t = 0;
  for (;;)
  {
    __syncthreads();
    if (threadIdx.x == 0)
    {
      int2 jbv = make_int2(-1, 0);
      while (t < nq)
      {
        const int q = (home + t) % nq;
        const unsigned n = static_cast<unsigned>(qoff[q + 1] - qoff[q]);
        const unsigned i = atomicAdd(&queue[q], 1u);
        if (i < n)
        {
          jbv = slots[qoff[q] + i];
          break;
        }
        ++t;
      }
      job = jbv;
    }
    __syncthreads();
    const int2 jb = job;
    if (jb.x < 0)
      break;
    const int32_t s = jb.x, m = jb.y, W = team[s];
    const int32_t r0s = sys_row[s], r1s = sys_row[s + 1];
    const int64_t ns = r1s - r0s;
    // this member's rows
    const int32_t lo = r0s + static_cast<int32_t>(ns * m / W);
    const int32_t hi = r0s + static_cast<int32_t>(ns * (m + 1) / W);
    const int g = gsel[s];
    unsigned* bar = bars + s;
    float* tred = tred_all + tred_off[s];
    // double sums after the float ones (an even float offset: 8-byte aligned)
    double* tredd = reinterpret_cast<double*>(tred + 2 * static_cast<int64_t>(W) * NRHS);
    unsigned nbar = 0;
    int nsum_d = 0;

    // x = 0, r = b, p = b, rr0 = ||b||^2 (cg.cpp:31-44); with a reordered
    // batch, row i of the solver is row perm[i] of b and x
    float* __restrict__ x = perm ? xin : xu;
    VecN<NRHS> rr{};
    for (int32_t i = lo + static_cast<int32_t>(threadIdx.x); i < hi; i += blockDim.x)
    {
      const int64_t o = static_cast<int64_t>(i) * NRHS;
      const VecN<NRHS> bi = ldn<NRHS>(b + static_cast<int64_t>(perm ? perm[i] : i) * NRHS);
      stn<NRHS>(x + o, VecN<NRHS>{});
      stn<NRHS>(r + o, bi);
      stn<NRHS>(p + o, bi);
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
        rr.v[j] += bi.v[j] * bi.v[j];
    }
    // ||b||^2 in double as well, for the true-residual check
    VecD<NRHS> bb{};
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      bb.v[j] = static_cast<double>(rr.v[j]);
    rr = team_sum<NRHS>(rr, tred, 0, W, m, bar, nbar, err, red); // also publishes p
    {
      VecD<NRHS> b2{};
      for (int32_t i = lo + static_cast<int32_t>(threadIdx.x); i < hi; i += blockDim.x)
      {
        const VecN<NRHS> bi = ldn<NRHS>(b + static_cast<int64_t>(perm ? perm[i] : i) * NRHS);
#pragma unroll
        for (int j = 0; j < NRHS; ++j)
          b2.v[j] += static_cast<double>(bi.v[j]) * static_cast<double>(bi.v[j]);
      }
      bb = team_sum_d<NRHS>(b2, tredd, (nsum_d++) & 1, W, m, bar, nbar, err, redd);
    }
    const VecN<NRHS> rr0 = rr;
    // per right-hand side stopping threshold on ||r||^2 / ||b||^2: rtol^2,
    // a quarter of it again after every failed true-residual check
    float tgt[NRHS];
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
      tgt[j] = rtol2;
    bool active[NRHS], by_rtol[NRHS];
    int32_t its[NRHS];
    bool any = false;
    int restarts = 0;
#pragma unroll
    for (int j = 0; j < NRHS; ++j)
    {
      active[j] = max_its > 0;
      by_rtol[j] = false;
      its[j] = 0;
      any = any || active[j];
    }
    for (int32_t k = 1; k <= max_its && any; ++k)
    {
      __syncthreads(); // p complete (within the workgroup; across the team: barrier below)
      VecN<NRHS> pdy;
      switch (g)
      {
      case 1: pdy = batch_spmv<NRHS, 1>(rowptr, colidx, val, p, y, lo, hi); break;
      case 2: pdy = batch_spmv<NRHS, 2>(rowptr, colidx, val, p, y, lo, hi); break;
      case 4: pdy = batch_spmv<NRHS, 4>(rowptr, colidx, val, p, y, lo, hi); break;
      case 8: pdy = batch_spmv<NRHS, 8>(rowptr, colidx, val, p, y, lo, hi); break;
      default: pdy = batch_spmv<NRHS, 16>(rowptr, colidx, val, p, y, lo, hi); break;
      }
      pdy = team_sum<NRHS>(pdy, tred, 1, W, m, bar, nbar, err, red);
      float alpha[NRHS];
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
        alpha[j] = active[j] ? rr.v[j] / pdy.v[j] : 0.0f;
      // r -= alpha y and ||r||^2; x += alpha p is applied in the pass that
      // forms the next direction (it reads p there anyway): one vector pass
      // fewer per iteration, the same operations on every element
      VecN<NRHS> rn{};
      for (int32_t i = lo + static_cast<int32_t>(threadIdx.x); i < hi; i += blockDim.x)
      {
        const int64_t o = static_cast<int64_t>(i) * NRHS;
        VecN<NRHS> ri = ldn<NRHS>(r + o);
        const VecN<NRHS> yi = ldn<NRHS>(y + o);
#pragma unroll
        for (int j = 0; j < NRHS; ++j)
        {
          ri.v[j] = ri.v[j] - alpha[j] * yi.v[j];
          rn.v[j] += ri.v[j] * ri.v[j];
        }
        stn<NRHS>(r + o, ri);
      }
      rn = team_sum<NRHS>(rn, tred, 0, W, m, bar, nbar, err, red);
      float beta[NRHS];
      any = false;
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
      {
        beta[j] = active[j] ? rn.v[j] / rr.v[j] : 0.0f;
        if (active[j])
        {
          its[j] = k;
          rr.v[j] = rn.v[j];
          if (rn.v[j] / rr0.v[j] < tgt[j])
          {
            active[j] = false;
            by_rtol[j] = true;
          }
          else if (!(rn.v[j] == rn.v[j]))
            active[j] = false;
        }
        any = any || active[j];
      }
      if (!any)
      {
        // last iteration: x += alpha p only
        for (int32_t i = lo + static_cast<int32_t>(threadIdx.x); i < hi; i += blockDim.x)
        {
          const int64_t o = static_cast<int64_t>(i) * NRHS;
          VecN<NRHS> xi = ldn<NRHS>(x + o);
          const VecN<NRHS> pi = ldn<NRHS>(p + o);
#pragma unroll
          for (int j = 0; j < NRHS; ++j)
            xi.v[j] = xi.v[j] + alpha[j] * pi.v[j];
          stn<NRHS>(x + o, xi);
        }
        if (restarts >= kMaxRestarts || k == max_its)
          break;
        // True-residual check of the right-hand sides that met their
        // threshold by the recursive residual: t = b - A x in double; where
        // ||t|| does not meet rtol, CG restarts from x with r = p = t and a
        // threshold four times lower (the recursive residual of FP32 CG
        // drifts from the true one by ~1e-7 relative).
        if (W > 1)
          team_barrier(bar, static_cast<unsigned>(W) * (++nbar), err); // x complete
        else
          __syncthreads();
        VecD<NRHS> tt;
        switch (g)
        {
        case 1: tt = batch_resid<NRHS, 1>(rowptr, colidx, val, perm, b, x, r, p, lo, hi); break;
        case 2: tt = batch_resid<NRHS, 2>(rowptr, colidx, val, perm, b, x, r, p, lo, hi); break;
        case 4: tt = batch_resid<NRHS, 4>(rowptr, colidx, val, perm, b, x, r, p, lo, hi); break;
        case 8: tt = batch_resid<NRHS, 8>(rowptr, colidx, val, perm, b, x, r, p, lo, hi); break;
        default: tt = batch_resid<NRHS, 16>(rowptr, colidx, val, perm, b, x, r, p, lo, hi); break;
        }
        // (also publishes r and p to the team)
        tt = team_sum_d<NRHS>(tt, tredd, (nsum_d++) & 1, W, m, bar, nbar, err, redd);
        ++restarts;
#pragma unroll
        for (int j = 0; j < NRHS; ++j)
        {
          if (by_rtol[j] && !(tt.v[j] < rtol2d * bb.v[j]))
          {
            active[j] = true;
            rr.v[j] = static_cast<float>(tt.v[j]);
            tgt[j] *= 0.25f;
            any = true;
          }
          by_rtol[j] = false;
        }
        if (!any)
          break;
        continue; // the next iteration's product reads p = r
      }
      for (int32_t i = lo + static_cast<int32_t>(threadIdx.x); i < hi; i += blockDim.x)
      {
        const int64_t o = static_cast<int64_t>(i) * NRHS;
        const VecN<NRHS> ri = ldn<NRHS>(r + o);
        VecN<NRHS> pi = ldn<NRHS>(p + o), xi = ldn<NRHS>(x + o);
#pragma unroll
        for (int j = 0; j < NRHS; ++j)
        {
          xi.v[j] = xi.v[j] + alpha[j] * pi.v[j];
          pi.v[j] = pi.v[j] * beta[j] + ri.v[j];
        }
        stn<NRHS>(x + o, xi);
        stn<NRHS>(p + o, pi);
      }
      if (W > 1)
        team_barrier(bar, static_cast<unsigned>(W) * (++nbar), err); // p complete
    }
    if (perm)
    {
      // this member's rows of x back to the caller's order
      __syncthreads();
      for (int32_t i = lo + static_cast<int32_t>(threadIdx.x); i < hi; i += blockDim.x)
        stn<NRHS>(xu + static_cast<int64_t>(perm[i]) * NRHS,
                  ldn<NRHS>(x + static_cast<int64_t>(i) * NRHS));
    }
    if (m == 0 && threadIdx.x == 0)
#pragma unroll
      for (int j = 0; j < NRHS; ++j)
        its_out[static_cast<int64_t>(s) * NRHS + j] = its[j];
  }
}

// queues = 0, every team's barrier counter = 0, err = 0 (a kernel, not a
// memset node, so that graph replays are reliable)
__global__ void batch_reset(unsigned* queue, int32_t nq, unsigned* bars, int32_t nsys, int* err)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nq)
    queue[i] = 0u;
  if (i == 0)
    *err = 0;
  for (int32_t k = i; k < nsys; k += gridDim.x * blockDim.x)
    bars[k] = 0u;
}

int pow2_ceil(int v)
{
  int g = 1;
  while (g < v)
    g <<= 1;
  return g;
}

} // namespace
} // namespace spmvx

using namespace spmvx;

struct spmvx_batch_s
{
  int32_t nsys = 0, nrhs = 0;
  int64_t nrows = 0, nnz = 0;
  std::vector<int32_t> h_sys_row, h_rowptr;
  DevBuf rowptr, colidx, val, sys_row, gsel, scratch, its, queue;
  // reverse Cuthill-McKee row order of every system (new row -> caller's row)
  DevBuf perm;
  bool reordered = false;
  // team plan
  DevBuf slots, qoff, team, tred_off, bars, tred, err;
  int32_t nslots = 0;
  int32_t nq = kNumXcd; // work queues, one per XCD
  bool plan_dirty = true;
  // launch configuration (autotuned)
  int32_t block = 1024;
  int32_t g_policy = 0;          // 0 = per system from the mean row length, else fixed G
  int32_t rows_per_member = 16384; // team size = ceil(rows / rows_per_member), 0 = no teams
  int grid = 0;
  // captured graph and the arguments it was captured with
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  const void* cap_b = nullptr;
  void* cap_x = nullptr;
  int32_t cap_its = -1;
  double cap_rtol = -1.0;
  // the batch's own stream (graphs cannot be captured on the null stream),
  // ordered after / before the caller's stream with events
  hipStream_t own = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  ~spmvx_batch_s()
  {
    if (exec)
      (void)hipGraphExecDestroy(exec);
    if (graph)
      (void)hipGraphDestroy(graph);
    if (ev_in)
      (void)hipEventDestroy(ev_in);
    if (ev_out)
      (void)hipEventDestroy(ev_out);
    if (own)
      (void)hipStreamDestroy(own);
  }
};

namespace
{

// Reverse Cuthill-McKee order of the rows of one system (George & Liu):
// breadth-first from a pseudo-peripheral row, the unvisited neighbours of
// each row taken in increasing degree, every connected component in turn,
// the whole order reversed.  rp/ci describe the system's rows with columns
// local to the system.  Rows that share columns end up close together, so
// the direction-vector gathers of neighbouring rows share cache lines.
std::vector<int32_t> rcm_order(int32_t n, const int32_t* rp, const int32_t* ci)
{
  std::vector<int32_t> order;
  order.reserve(static_cast<size_t>(n));
  std::vector<int32_t> stamp(static_cast<size_t>(n), -1), vis, nb;
  std::vector<char> done(static_cast<size_t>(n), 0);
  int32_t cur = 0;
  auto deg = [&](int32_t u) { return rp[u + 1] - rp[u]; };
  // breadth-first levels from s; returns the depth and the last level's start
  auto bfs = [&](int32_t s0, size_t& last) {
    ++cur;
    vis.assign(1, s0);
    stamp[s0] = cur;
    size_t begin = 0, end = 1;
    int depth = 0;
    for (;;)
    {
      for (size_t h = begin; h < end; ++h)
        for (int32_t k = rp[vis[h]]; k < rp[vis[h] + 1]; ++k)
          if (stamp[ci[k]] != cur)
          {
            stamp[ci[k]] = cur;
            vis.push_back(ci[k]);
          }
      if (vis.size() == end)
      {
        last = begin;
        return depth;
      }
      begin = end;
      end = vis.size();
      ++depth;
    }
  };
  for (int32_t start = 0; start < n; ++start)
  {
    if (done[start])
      continue;
    // pseudo-peripheral row of this component
    int32_t s0 = start;
    size_t last = 0;
    int d = bfs(s0, last);
    for (int t = 0; t < 4; ++t)
    {
      int32_t best = vis[last];
      for (size_t h = last; h < vis.size(); ++h)
        if (deg(vis[h]) < deg(best))
          best = vis[h];
      size_t last2 = 0;
      const int d2 = bfs(best, last2);
      if (d2 <= d)
        break;
      s0 = best;
      d = d2;
      last = last2;
    }
    // Cuthill-McKee from s0
    size_t head = order.size();
    order.push_back(s0);
    done[s0] = 1;
    while (head < order.size())
    {
      const int32_t u = order[head++];
      nb.clear();
      for (int32_t k = rp[u]; k < rp[u + 1]; ++k)
        if (!done[ci[k]])
        {
          done[ci[k]] = 1;
          nb.push_back(ci[k]);
        }
      std::stable_sort(nb.begin(), nb.end(),
                       [&](int32_t a, int32_t c) { return deg(a) < deg(c); });
      order.insert(order.end(), nb.begin(), nb.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

int batch_capacity(spmvx_batch_s* B)
{
  int per_cu = 0;
  if (B->nrhs == 4)
    SPMVX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_cg_kernel<4>, B->block, 0));
  else if (B->nrhs == 2)
    SPMVX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_cg_kernel<2>, B->block, 0));
  else
    SPMVX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_cg_kernel<1>, B->block, 0));
  int dev = 0, ncu = 0;
  SPMVX_HIP(hipGetDevice(&dev));
  SPMVX_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(1, per_cu) * std::max(1, ncu);
}

// Lanes per row (per system) and the team plan: systems in decreasing size
// order (LPT), each split over W workgroups (W <= capacity / 2 so that a team
// can always be completed by running workgroups), one queue slot per member.
void build_plan(spmvx_batch_s* B, int capacity)
{
  const int nsys = B->nsys;
  std::vector<int8_t> g(nsys);
  std::vector<int32_t> team(nsys);
  std::vector<int64_t> work(nsys);
  // every XCD queue may hold one team still gathering members; capping teams
  // at capacity / (2 * queues) leaves enough running workgroups to complete
  // all of them
  const int wmax = std::max(1, std::min(64, capacity / (2 * B->nq)));
  for (int s = 0; s < nsys; ++s)
  {
    const int64_t rows = B->h_sys_row[s + 1] - B->h_sys_row[s];
    const int64_t nz = B->h_rowptr[B->h_sys_row[s + 1]] - B->h_rowptr[B->h_sys_row[s]];
    work[s] = rows + nz;
    if (B->g_policy > 0)
      g[s] = static_cast<int8_t>(B->g_policy);
    else
    {
      const int mean = rows > 0 ? static_cast<int>((nz + rows - 1) / rows) : 1;
      g[s] = static_cast<int8_t>(std::min(16, pow2_ceil(std::max(1, mean / 4))));
    }
    int64_t w = 1;
    if (B->rows_per_member > 0)
      w = (rows + B->rows_per_member - 1) / B->rows_per_member;
    team[s] = static_cast<int32_t>(std::max<int64_t>(1, std::min<int64_t>(w, wmax)));
  }
  // LPT: systems by decreasing work go to the XCD queue with the least work
  std::vector<int32_t> order(nsys);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t a, int32_t c) { return work[a] > work[c]; });
  const int nq = B->nq;
  std::vector<std::vector<int2>> qs(nq);
  std::vector<int64_t> load(nq, 0);
  for (int32_t s : order)
  {
    const int q = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
    load[q] += work[s];
    for (int32_t m = 0; m < team[s]; ++m)
      qs[q].push_back(make_int2(s, m));
  }
  std::vector<int2> slots;
  std::vector<int32_t> qoff(nq + 1, 0);
  for (int q = 0; q < nq; ++q)
  {
    slots.insert(slots.end(), qs[q].begin(), qs[q].end());
    qoff[q + 1] = static_cast<int32_t>(slots.size());
  }
  std::vector<int64_t> off(nsys + 1, 0);
  for (int s = 0; s < nsys; ++s)
    off[s + 1] = off[s] + 6LL * team[s] * B->nrhs; // 2 float + 2 double buffers
  B->nslots = static_cast<int32_t>(slots.size());
  B->gsel.ensure(std::max(nsys, 1));
  B->team.ensure(4 * std::max(nsys, 1));
  B->tred_off.ensure(8 * std::max(nsys, 1));
  B->slots.ensure(sizeof(int2) * std::max<size_t>(slots.size(), 1));
  B->qoff.ensure(4 * (nq + 1));
  SPMVX_HIP(hipMemcpy(B->qoff.ptr, qoff.data(), 4 * (nq + 1), hipMemcpyHostToDevice));
  B->tred.ensure(4 * std::max<int64_t>(off[nsys], 1));
  B->bars.ensure(4 * std::max(nsys, 1));
  B->err.ensure(256);
  if (nsys)
  {
    SPMVX_HIP(hipMemcpy(B->gsel.ptr, g.data(), nsys, hipMemcpyHostToDevice));
    SPMVX_HIP(hipMemcpy(B->team.ptr, team.data(), 4 * nsys, hipMemcpyHostToDevice));
    SPMVX_HIP(hipMemcpy(B->tred_off.ptr, off.data(), 8 * nsys, hipMemcpyHostToDevice));
    SPMVX_HIP(hipMemcpy(B->slots.ptr, slots.data(), sizeof(int2) * slots.size(), hipMemcpyHostToDevice));
  }
  B->plan_dirty = false;
}

void launch_batch(spmvx_batch_s* B, const float* b, float* x, int32_t max_its, double rtol,
                  hipStream_t s)
{
  const int64_t vn = B->nrows * B->nrhs;
  float* r = B->scratch.as<float>();
  float* p = r + vn;
  float* y = p + vn;
  float* xin = y + vn; // x in the solver's row order (reordered batches)
  const int32_t* perm = B->reordered ? B->perm.as<int32_t>() : nullptr;
  batch_reset<<<std::max(1, std::min(64, (B->nsys + 255) / 256)), 256, 0, s>>>(
      B->queue.as<unsigned>(), B->nq, B->bars.as<unsigned>(), B->nsys, B->err.as<int>());
#define SPMVX_BATCH(NR)                                                                        \
  batch_cg_kernel<NR><<<B->grid, B->block, 0, s>>>(                                           \
      B->rowptr.as<int32_t>(), B->colidx.as<int32_t>(), B->val.as<float>(),                    \
      B->sys_row.as<int32_t>(), B->slots.as<int2>(), B->qoff.as<int32_t>(), B->nq,            \
      B->team.as<int32_t>(),                                                                   \
      B->tred_off.as<int64_t>(), B->gsel.as<int8_t>(), perm, b, x, xin, r, p, y, max_its,      \
      static_cast<float>(rtol * rtol), rtol * rtol,                                            \
      B->its.as<int32_t>(), B->queue.as<unsigned>(), B->bars.as<unsigned>(),                   \
      B->tred.as<float>(), B->err.as<int>())
  if (B->nrhs == 4)
    SPMVX_BATCH(4);
  else if (B->nrhs == 2)
    SPMVX_BATCH(2);
  else
    SPMVX_BATCH(1);
#undef SPMVX_BATCH
}

void drop_graph(spmvx_batch_s* B)
{
  if (B->exec)
    SPMVX_HIP(hipGraphExecDestroy(B->exec));
  if (B->graph)
    SPMVX_HIP(hipGraphDestroy(B->graph));
  B->exec = nullptr;
  B->graph = nullptr;
}

} // namespace

extern "C" {

int spmvx_batch_create(int32_t nsys, const int32_t* sys_row, const int32_t* rowptr,
                       const int32_t* colidx, const float* values, int32_t nrhs,
                       spmvx_batch* out)
{
  return guarded([&] {
    SPMVX_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    SPMVX_REQUIRE(nsys >= 0 && sys_row != nullptr && rowptr != nullptr, "invalid batch");
    SPMVX_REQUIRE(nrhs == 1 || nrhs == 2 || nrhs == 4, "nrhs must be 1, 2 or 4");
    SPMVX_REQUIRE(sys_row[0] == 0, "sys_row[0] must be 0");
    for (int s = 0; s < nsys; ++s)
      SPMVX_REQUIRE(sys_row[s + 1] >= sys_row[s], "sys_row must be non-decreasing");
    const int64_t nrows = sys_row[nsys];
    const int64_t nnz = rowptr[nrows];
    SPMVX_REQUIRE(rowptr[0] == 0, "rowptr[0] must be 0");
    for (int64_t i = 0; i < nrows; ++i)
      SPMVX_REQUIRE(rowptr[i + 1] >= rowptr[i], "rowptr must be non-decreasing");
    SPMVX_REQUIRE(nnz == 0 || (colidx && values), "NULL arrays");
    for (int s = 0; s < nsys; ++s)
      for (int64_t i = sys_row[s]; i < sys_row[s + 1]; ++i)
        for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
          SPMVX_REQUIRE(colidx[k] >= sys_row[s] && colidx[k] < sys_row[s + 1],
                        "column outside its system (matrix must be block diagonal)");
    auto B = std::make_unique<spmvx_batch_s>();
    B->nsys = nsys;
    B->nrhs = nrhs;
    B->nrows = nrows;
    B->nnz = nnz;
    B->h_sys_row.assign(sys_row, sys_row + nsys + 1);
    // SPMVX_BATCH_REORDER=1: every system reordered by reverse Cuthill-McKee
    // (rows, columns and values permuted here, b and x read and written
    // through the permutation by the solver).  Off by default: on the
    // uniformly random symmetric patterns of the benchmark it changed the
    // solve time by under 1% (the p gathers stay scattered).
    const char* env = std::getenv("SPMVX_BATCH_REORDER");
    B->reordered = env && std::atoi(env) == 1 && nrows > 0;
    std::vector<int32_t> nrp, nci, perm;
    std::vector<float> nva;
    if (B->reordered)
    {
      nrp.assign(static_cast<size_t>(nrows) + 1, 0);
      nci.resize(static_cast<size_t>(nnz));
      nva.resize(static_cast<size_t>(nnz));
      perm.resize(static_cast<size_t>(nrows));
      std::vector<int32_t> inv(static_cast<size_t>(nrows));
      auto work = [&](int s) {
        const int32_t base = sys_row[s], n = sys_row[s + 1] - base;
        std::vector<int32_t> lrp(static_cast<size_t>(n) + 1), lci;
        const int32_t e0 = rowptr[base];
        lci.resize(static_cast<size_t>(rowptr[base + n] - e0));
        for (int32_t i = 0; i <= n; ++i)
          lrp[i] = rowptr[base + i] - e0;
        for (size_t k = 0; k < lci.size(); ++k)
          lci[k] = colidx[e0 + k] - base;
        const std::vector<int32_t> ord = rcm_order(n, lrp.data(), lci.data());
        for (int32_t i = 0; i < n; ++i)
        {
          perm[base + i] = base + ord[i];
          inv[base + ord[i]] = base + i;
        }
      };
      const int nth = std::max(1, std::min<int>(16, static_cast<int>(std::thread::hardware_concurrency())));
      {
        std::vector<std::thread> pool;
        std::atomic<int> next{0};
        for (int t = 0; t < nth; ++t)
          pool.emplace_back([&] {
            for (int s; (s = next.fetch_add(1)) < nsys;)
              work(s);
          });
        for (auto& th : pool)
          th.join();
      }
      for (int64_t i = 0; i < nrows; ++i)
        nrp[i + 1] = nrp[i] + (rowptr[perm[i] + 1] - rowptr[perm[i]]);
      auto fill = [&](int64_t i0, int64_t i1) {
        std::vector<std::pair<int32_t, float>> row;
        for (int64_t i = i0; i < i1; ++i)
        {
          row.clear();
          for (int32_t k = rowptr[perm[i]]; k < rowptr[perm[i] + 1]; ++k)
            row.emplace_back(inv[colidx[k]], values[k]);
          std::sort(row.begin(), row.end(),
                    [](const std::pair<int32_t, float>& a, const std::pair<int32_t, float>& c) {
                      return a.first < c.first;
                    });
          for (size_t q = 0; q < row.size(); ++q)
          {
            nci[nrp[i] + q] = row[q].first;
            nva[nrp[i] + q] = row[q].second;
          }
        }
      };
      {
        std::vector<std::thread> pool;
        for (int t = 0; t < nth; ++t)
          pool.emplace_back(fill, nrows * t / nth, nrows * (t + 1) / nth);
        for (auto& th : pool)
          th.join();
      }
      rowptr = nrp.data();
      colidx = nci.data();
      values = nva.data();
      B->perm.ensure(4 * static_cast<size_t>(nrows));
      SPMVX_HIP(hipMemcpy(B->perm.ptr, perm.data(), 4 * nrows, hipMemcpyHostToDevice));
    }
    B->h_rowptr.assign(rowptr, rowptr + nrows + 1);
    B->rowptr.ensure(4 * (nrows + 1));
    SPMVX_HIP(hipMemcpy(B->rowptr.ptr, rowptr, 4 * (nrows + 1), hipMemcpyHostToDevice));
    B->colidx.ensure(4 * std::max<int64_t>(nnz, 1));
    B->val.ensure(4 * std::max<int64_t>(nnz, 1));
    if (nnz)
    {
      SPMVX_HIP(hipMemcpy(B->colidx.ptr, colidx, 4 * nnz, hipMemcpyHostToDevice));
      SPMVX_HIP(hipMemcpy(B->val.ptr, values, 4 * nnz, hipMemcpyHostToDevice));
    }
    B->sys_row.ensure(4 * (nsys + 1));
    SPMVX_HIP(hipMemcpy(B->sys_row.ptr, sys_row, 4 * (nsys + 1), hipMemcpyHostToDevice));
    // r, p, y and x in the solver's row order
    B->scratch.ensure(sizeof(float) * 4 * static_cast<size_t>(std::max<int64_t>(nrows, 1)) * nrhs);
    B->its.ensure(4 * static_cast<size_t>(std::max(nsys, 1)) * nrhs);
    B->queue.ensure(256);
    *out = B.release();
  });
}

int spmvx_batch_destroy(spmvx_batch B)
{
  return guarded([&] { delete B; });
}

int spmvx_batch_set_config(spmvx_batch B, int32_t block, int32_t lanes_per_row,
                           int32_t rows_per_member)
{
  return guarded([&] {
    SPMVX_REQUIRE(B != nullptr, "NULL batch");
    SPMVX_REQUIRE(block == 256 || block == 512 || block == 1024, "block must be 256, 512 or 1024");
    SPMVX_REQUIRE(lanes_per_row == 0 || lanes_per_row == 1 || lanes_per_row == 2
                      || lanes_per_row == 4 || lanes_per_row == 8 || lanes_per_row == 16,
                  "lanes_per_row must be 0 (per system), 1, 2, 4, 8 or 16");
    SPMVX_REQUIRE(rows_per_member >= 0, "rows_per_member must be >= 0");
    if (block != B->block || lanes_per_row != B->g_policy || rows_per_member != B->rows_per_member)
    {
      B->block = block;
      B->g_policy = lanes_per_row;
      B->rows_per_member = rows_per_member;
      B->plan_dirty = true;
      drop_graph(B);
    }
  });
}

int spmvx_batch_solve(spmvx_batch B, const float* b, float* x, int32_t max_its, double rtol,
                      int32_t* its_host, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(B != nullptr, "NULL batch");
    SPMVX_REQUIRE(max_its >= 0, "max_its must be >= 0");
    SPMVX_REQUIRE(B->nrows == 0 || (b && x), "NULL vector");
    hipStream_t s = as_stream(stream);
    if (B->nsys == 0)
      return;
    if (B->plan_dirty)
    {
      const int cap = batch_capacity(B);
      build_plan(B, cap);
      // persistent grid: every workgroup resident at once
      B->grid = std::max(1, std::min(B->nslots, cap));
      drop_graph(B);
    }
    if (!B->own)
    {
      SPMVX_HIP(hipStreamCreateWithFlags(&B->own, hipStreamNonBlocking));
      SPMVX_HIP(hipEventCreateWithFlags(&B->ev_in, hipEventDisableTiming));
      SPMVX_HIP(hipEventCreateWithFlags(&B->ev_out, hipEventDisableTiming));
    }
    if (!B->exec || B->cap_b != b || B->cap_x != x || B->cap_its != max_its
        || B->cap_rtol != rtol)
    {
      drop_graph(B);
      B->graph = capture_graph(B->own, [&] { launch_batch(B, b, x, max_its, rtol, B->own); });
      SPMVX_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
      B->cap_b = b;
      B->cap_x = x;
      B->cap_its = max_its;
      B->cap_rtol = rtol;
    }
    SPMVX_HIP(hipEventRecord(B->ev_in, s));
    SPMVX_HIP(hipStreamWaitEvent(B->own, B->ev_in, 0));
    SPMVX_HIP(hipGraphLaunch(B->exec, B->own));
    SPMVX_HIP(hipEventRecord(B->ev_out, B->own));
    SPMVX_HIP(hipStreamWaitEvent(s, B->ev_out, 0));
    // the failure flag is always read (a timed-out team wait leaves invalid
    // results): the call synchronises the stream
    int err = 0;
    if (its_host)
      SPMVX_HIP(hipMemcpyAsync(its_host, B->its.ptr, 4 * static_cast<size_t>(B->nsys) * B->nrhs,
                               hipMemcpyDeviceToHost, s));
    SPMVX_HIP(hipMemcpyAsync(&err, B->err.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    SPMVX_HIP(hipStreamSynchronize(s));
    if (err)
      throw Error(SPMVX_ERR_STATE, "batched CG: a team barrier timed out (results invalid)");
  });
}

int spmvx_batch_info(spmvx_batch B, int32_t* nsys, int64_t* nrows, int64_t* nnz, int32_t* nrhs,
                     int32_t* block, int32_t* lanes_per_row, int32_t* rows_per_member,
                     int32_t* grid, int32_t* nslots)
{
  return guarded([&] {
    SPMVX_REQUIRE(B != nullptr, "NULL batch");
    if (nsys) *nsys = B->nsys;
    if (nrows) *nrows = B->nrows;
    if (nnz) *nnz = B->nnz;
    if (nrhs) *nrhs = B->nrhs;
    if (block) *block = B->block;
    if (lanes_per_row) *lanes_per_row = B->g_policy;
    if (rows_per_member) *rows_per_member = B->rows_per_member;
    if (grid) *grid = B->grid;
    if (nslots) *nslots = B->nslots;
  });
}

} // extern "C"
