// Device kernels of libspmv_amd (gfx950).
//
// CSR SpMV, "ordered stream" kernel
// ---------------------------------
// The analysis (matrix.hip) cuts the rows into blocks of at most kBS rows
// holding at most CHUNK stored entries (or one longer row).  One workgroup
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

#include <initializer_list>

#include <hip/hip_ext.h>

#include "common.hpp"

namespace spmvx
{

constexpr int kBS = 256; // threads per SpMV workgroup = max rows per block

// ---------------------------------------------------------------- values
struct bf16_t
{
  uint16_t bits;
};

// Complex values (Matrix<std::complex<T>>, Matrix.cpp:441-444): the product
// and sum std::complex evaluates without FMA, re = ar xr - ai xi,
// im = ar xi + ai xr, every product and sum rounded separately.
template <typename T>
struct cplx
{
  T re, im;
  cplx() = default;
  __host__ __device__ constexpr cplx(T r, T i = T(0)) : re(r), im(i) {}
};
template <typename T>
__host__ __device__ inline cplx<T> operator*(cplx<T> a, cplx<T> b)
{
  return cplx<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__host__ __device__ inline cplx<T> operator+(cplx<T> a, cplx<T> b)
{
  return cplx<T>(a.re + b.re, a.im + b.im);
}
template <typename T>
__host__ __device__ inline cplx<T> operator-(cplx<T> a, cplx<T> b)
{
  return cplx<T>(a.re - b.re, a.im - b.im);
}
template <typename T>
__host__ __device__ inline cplx<T>& operator+=(cplx<T>& a, cplx<T> b)
{
  a.re += b.re;
  a.im += b.im;
  return a;
}
template <typename T>
constexpr bool kIsComplex = false;
template <typename T>
constexpr bool kIsComplex<cplx<T>> = true;


// ---------------------------------------------------------------- CG state
// Scalars of one CG solve, resident on the device.  Every kernel of an
// iteration checks the state at its start and returns at once when the solve
// has stopped, so the host (or a replayed graph of a fixed number of
// iterations) can enqueue more iterations than are needed and the iteration
// count is still exact.  No kernel synchronises with another inside a launch:
// reductions write per-block partial sums and a one-workgroup finalize
// kernel (after the kernel boundary) combines them in a fixed order.
template <typename Acc>
struct CgState
{
  Acc rnorm0;
  Acc rnorm_old;
  Acc alpha;
  Acc beta;
  Acc rtol2;
  Acc last;           // last reduction result (residual kernels)
  int32_t k;          // iterations completed
  int32_t kmax;
  int32_t stopped;    // converged or k == kmax
  int32_t converged;
  int32_t xp_pending; // the residual update of iteration k ran; x/p update due
};

// Kernels before the residual update run iff this holds (live = 0 means
// "not part of a CG iteration": always run).
template <typename Acc>
__device__ __forceinline__ bool cg_live(const CgState<Acc>* st, int32_t live)
{
  return live == 0 || !st->stopped;
}

enum SpmvMode
{
  kPlain = 0, // y = A x
  kDot = 1,   // y = A p, partial p.y                      (CG, cg.cpp:56-61)
  kResid = 2  // y = b - A x, partial ||y||^2               (refinement)
};

// Sum of n per-block partials by one workgroup (blockDim.x = kFinBS), fixed
// order.  Result valid in all threads.
constexpr int kFinBS = 1024;
template <typename Acc>
__device__ __forceinline__ Acc finalize_sum(const Acc* __restrict__ partials, int64_t n,
                                            Acc* red)
{
  // 16 independent loads per thread and step, the last step predicated
  // (latency-bound: tens of thousands of partials take one to four round
  // trips), four accumulators, fixed order
  Acc s0 = Acc(0), s1 = Acc(0), s2 = Acc(0), s3 = Acc(0);
  for (int64_t i = threadIdx.x; i < n; i += 16 * kFinBS)
  {
    Acc v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u)
    {
      const int64_t j = i + u * kFinBS;
      v[u] = j < n ? partials[j] : Acc(0);
    }
#pragma unroll
    for (int u = 0; u < 16; u += 4)
    {
      s0 += v[u];
      s1 += v[u + 1];
      s2 += v[u + 2];
      s3 += v[u + 3];
    }
  }
  return block_sum((s0 + s1) + (s2 + s3), red);
}

// Streaming loads of the matrix arrays (read once per product): non-temporal
// loads keep them from displacing x, which is re-read across neighbouring
// row blocks, from L2.
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T* p)
{
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}
template <bool NT>
__device__ __forceinline__ double ld_val(const double* v, int64_t k) { return ld_stream<NT>(v + k); }
template <bool NT>
__device__ __forceinline__ float ld_val(const float* v, int64_t k) { return ld_stream<NT>(v + k); }
template <bool NT, typename T>
__device__ __forceinline__ cplx<T> ld_val(const cplx<T>* v, int64_t k)
{
  return v[k];
}
template <bool NT>
__device__ __forceinline__ float ld_val(const bf16_t* v, int64_t k)
{
  const uint16_t b = ld_stream<NT>(reinterpret_cast<const uint16_t*>(v) + k);
  return __uint_as_float(static_cast<uint32_t>(b) << 16);
}

// Loads N consecutive elements (N * sizeof(T) bytes, a multiple of 4 and
// aligned to min(16, N * sizeof(T))) as 16/8/4-byte vector loads.
template <bool NT, typename T, int N>
__device__ __forceinline__ void load_vec(const T* p, T (&out)[N])
{
  constexpr int B = N * static_cast<int>(sizeof(T));
  static_assert(B % 4 == 0 || B == 2, "vector load of whole dwords");
  if constexpr (B >= 16)
  {
    typedef uint32_t u4 __attribute__((ext_vector_type(4)));
    const u4* q = reinterpret_cast<const u4*>(p);
#pragma unroll
    for (int i = 0; i < B / 16; ++i)
    {
      const u4 v = ld_stream<NT>(q + i);
      __builtin_memcpy(reinterpret_cast<char*>(out) + 16 * i, &v, 16);
    }
  }
  else if constexpr (B == 8)
  {
    typedef uint32_t u2 __attribute__((ext_vector_type(2)));
    const u2 v = ld_stream<NT>(reinterpret_cast<const u2*>(p));
    __builtin_memcpy(out, &v, 8);
  }
  else if constexpr (B == 4)
  {
    const uint32_t v = ld_stream<NT>(reinterpret_cast<const uint32_t*>(p));
    __builtin_memcpy(out, &v, 4);
  }
  else
  {
    const uint16_t v = ld_stream<NT>(reinterpret_cast<const uint16_t*>(p));
    __builtin_memcpy(out, &v, 2);
  }
}

// Result store of the SpMV kernels: a relaxed agent-scope atomic store, i.e.
// an ordinary store with the agent-coherent cache policy.  In a kernel that
// streams ~40x more bytes than it writes, plain stores of y cost ~2.7% more
// (measured on the 27-point share).
template <typename T>
__device__ __forceinline__ void store_result(T* p, T v)
{
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void store_result(cplx<T>* p, cplx<T> v)
{
  store_result(&p->re, v.re);
  store_result(&p->im, v.im);
}

template <typename V>
__device__ __forceinline__ auto widen(V v)
{
  if constexpr (sizeof(V) == 2)
    return __uint_as_float(static_cast<uint32_t>(reinterpret_cast<const uint16_t&>(v)) << 16);
  else
    return v;
}

// Sequential sum prod[a] + prod[a+1] + ... + prod[e-1] starting from 0 (the
// reference's per-row order), with the LDS reads issued four at a time.
template <typename Acc>
__device__ __forceinline__ Acc row_sum(const Acc* prod, int a, int e)
{
  Acc s = Acc(0);
  int q = a;
  for (; q + 4 <= e; q += 4)
  {
    const Acc p0 = prod[q], p1 = prod[q + 1], p2 = prod[q + 2], p3 = prod[q + 3];
    s += p0;
    s += p1;
    s += p2;
    s += p3;
  }
  for (; q < e; ++q)
    s += prod[q];
  return s;
}

// Sum of a row longer than a block's staging chunk (entries [k0, k1)),
// valid in thread 0.  Ordered (par == 0): the rounded products are staged
// CHUNK at a time in LDS and thread 0 adds them in stored order, starting from
// 0 -- the sequential product, bit-identical to the reference; its cost is
// one dependent add per entry in one lane.  Parallel (par != 0, real values):
// thread t adds the products of entries k0 + t, k0 + t + BS, ... in that
// order and the workgroup's partial sums are combined in a fixed tree
// (block_sum): deterministic, but rounded differently from the sequential
// sum.  Called by all threads; starts after and ends with a barrier.
template <typename V, typename X, typename Acc, int CHUNK, int BS, bool NT>
__device__ __forceinline__ Acc long_row_sum(const int32_t* __restrict__ colidx,
                                            const V* __restrict__ val, const X* __restrict__ x,
                                            int64_t k0, int64_t k1, Acc* prod, Acc* red,
                                            int32_t par)
{
  const int tid = threadIdx.x;
  if constexpr (!kIsComplex<Acc>)
  {
    if (par)
    {
      // eight entries' loads in flight per thread; the adds stay in order
      constexpr int U = 8;
      Acc s = Acc(0);
      int64_t k = k0 + tid;
      for (; k + (U - 1) * BS < k1; k += U * BS)
      {
        int32_t c[U];
        Acc a[U];
#pragma unroll
        for (int j = 0; j < U; ++j)
        {
          c[j] = ld_stream<NT>(colidx + k + j * BS);
          a[j] = static_cast<Acc>(ld_val<NT>(val, k + j * BS));
        }
        X xv[U];
#pragma unroll
        for (int j = 0; j < U; ++j)
          xv[j] = x[c[j]];
#pragma unroll
        for (int j = 0; j < U; ++j)
          s += a[j] * static_cast<Acc>(xv[j]);
      }
      for (; k < k1; k += BS)
        s += static_cast<Acc>(ld_val<NT>(val, k)) * static_cast<Acc>(x[ld_stream<NT>(colidx + k)]);
      s = block_sum<Acc, true>(s, red);
      lds_barrier(); // red is reused by the caller
      return s;
    }
  }
  Acc s = Acc(0);
  for (int64_t c0 = k0; c0 < k1; c0 += CHUNK)
  {
    const int n = static_cast<int>((k1 - c0) < CHUNK ? (k1 - c0) : CHUNK);
    for (int t = tid; t < n; t += BS)
    {
      const int64_t k = c0 + t;
      prod[t] = static_cast<Acc>(ld_val<NT>(val, k)) * static_cast<Acc>(x[colidx[k]]);
    }
    lds_barrier();
    if (tid == 0)
      for (int q = 0; q < n; ++q)
        s += prod[q];
    lds_barrier();
  }
  return s;
}

// Kernel configuration: products staged per block (CHUNK), consecutive
// entries each thread loads with one vector load per array (VEC; the entry
// range of a block is widened down to a multiple of VEC, the arrays are
// padded, and entries outside the block are skipped), independent vector
// steps in flight per thread (UNROLL), non-temporal matrix loads (NT).  Row
// blocks hold at most BS rows (the workgroup size) and CHUNK entries.  Block
// descriptors are
// (first row, first entry) pairs.
//
// Short column indices (IM != 0): when every row block's columns span less
// than 65536 (IM 1) the analysis stores each column as its offset from the
// block's first column (wbase.x); otherwise, when they fall into at most four
// windows of 16384 consecutive columns (IM 2: the planes of a 3D stencil),
// as (window << 14) | (offset in the window), with the four window bases in
// wbase.  Stencils, banded and locally numbered matrices fit.  The kernel
// decodes the same columns in the same order (bit-identical results) while
// streaming 2 instead of 4 bytes of index per entry.
template <typename V, typename X, typename Acc, int MODE, int CHUNK, int VEC, int UNROLL, bool NT,
          int BS, int IM>
__global__ __launch_bounds__(BS) void csr_stream(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const uint16_t* __restrict__ col16, const int4* __restrict__ wbase, int32_t ncols,
    const V* __restrict__ val, const X* __restrict__ x, X* __restrict__ y,
    const int2* __restrict__ blocks, int64_t nblocks, const X* __restrict__ b,
    Acc* __restrict__ partials, const CgState<Acc>* st, int32_t it, int32_t long_par)
{
  __shared__ Acc prod[CHUNK + 1];
  __shared__ Acc red[16];

  const int64_t bi = xcd_block(blockIdx.x, nblocks);
  const int2 d0 = blocks[bi], d1 = blocks[bi + 1];
  // A stopped CG iteration returns once both the state and the block
  // descriptor have arrived (the test reads both), so that the two loads are
  // in flight together rather than one after the other.
  if constexpr (MODE != kPlain)
    if (!cg_live(st, it) & (d1.y >= d0.y))
      return;
  const int32_t r0 = d0.x, r1 = d1.x;
  const int32_t k0 = d0.y, k1 = d1.y;
  const int tid = threadIdx.x;
  const int32_t row = r0 + tid;
  // row extents and the row's x (p.y) or b (residual) entry, loaded early
  // (independent of the staging below)
  int32_t ra = 0, re = 0;
  X xr = X(0);
  if (row < r1)
  {
    ra = rowptr[row];
    re = rowptr[row + 1];
    if constexpr (MODE == kDot)
      xr = x[row];
    else if constexpr (MODE == kResid)
      xr = b[row];
  }
  Acc part = Acc(0);
  int4 w = make_int4(0, 0, 0, 0); // base / window bases of a short-index block
  if constexpr (IM != 0)
    w = wbase[bi];

  if (k1 - k0 <= CHUNK)
  {
    const int32_t kb = k0 & ~(VEC - 1);
    constexpr int STEP = BS * VEC;
    for (int32_t g0 = kb; g0 < k1; g0 += STEP * UNROLL)
    {
      // Branch-free loads and gathers: a vector past the block is replaced by
      // the block's first vector, and entries outside the block (neighbouring
      // blocks' entries or the zeroed padding) are valid columns, so every
      // load can be issued before any result is needed; only the LDS store
      // is predicated.  (A short index decoded with this block's bases is
      // clamped into [0, ncols): inside the block it already is.)
      int32_t c[UNROLL][VEC];
      V v[UNROLL][VEC];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
      {
        const int32_t g = g0 + u * STEP + tid * VEC;
        const int32_t gs = g < k1 ? g : kb;
        if constexpr (IM != 0)
        {
          uint16_t e[VEC];
          load_vec<NT>(col16 + gs, e);
#pragma unroll
          for (int j = 0; j < VEC; ++j)
          {
            const uint32_t q = e[j];
            int32_t cc;
            if constexpr (IM == 1)
              cc = w.x + static_cast<int32_t>(q);
            else
            {
              const uint32_t sel = q >> 14;
              const int32_t wb = sel == 0 ? w.x : (sel == 1 ? w.y : (sel == 2 ? w.z : w.w));
              cc = wb + static_cast<int32_t>(q & 0x3FFFu);
            }
            c[u][j] = cc < ncols ? cc : ncols - 1;
          }
        }
        else
          load_vec<NT>(colidx + gs, c[u]);
        load_vec<NT>(val + gs, v[u]);
      }
      X xv[UNROLL][VEC];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
#pragma unroll
        for (int j = 0; j < VEC; ++j)
          xv[u][j] = x[c[u][j]];
#pragma unroll
      for (int u = 0; u < UNROLL; ++u)
      {
        const int32_t g = g0 + u * STEP + tid * VEC;
#pragma unroll
        for (int j = 0; j < VEC; ++j)
        {
          // entries outside the block go to the spare slot prod[CHUNK]: an
          // unconditional store keeps the compiler from sinking a gather
          // into a branch (where it would be issued after the others return)
          const int32_t k = g + j;
          const bool in = g < k1 && k >= k0 && k < k1;
          prod[in ? k - k0 : CHUNK] =
              static_cast<Acc>(widen(v[u][j])) * static_cast<Acc>(xv[u][j]);
        }
      }
    }
    lds_barrier();
    if (row < r1)
    {
      const Acc s = row_sum(prod, ra - k0, re - k0);
      if (MODE == kResid)
      {
        const X yi = xr - static_cast<X>(s);
        store_result(y + row, yi);
        part = static_cast<Acc>(yi) * static_cast<Acc>(yi);
      }
      else
      {
        const X yi = static_cast<X>(s);
        store_result(y + row, yi);
        if (MODE == kDot)
          part = static_cast<Acc>(xr) * static_cast<Acc>(yi);
      }
    }
  }
  else
  {
    // one row longer than CHUNK
    const Acc s =
        long_row_sum<V, X, Acc, CHUNK, BS, NT>(colidx, val, x, k0, k1, prod, red, long_par);
    if (tid == 0)
    {
      if (MODE == kResid)
      {
        const X yi = b[r0] - static_cast<X>(s);
        store_result(y + r0, yi);
        part = static_cast<Acc>(yi) * static_cast<Acc>(yi);
      }
      else
      {
        const X yi = static_cast<X>(s);
        store_result(y + r0, yi);
        if (MODE == kDot)
          part = static_cast<Acc>(x[r0]) * static_cast<Acc>(yi);
      }
    }
  }

  if constexpr (MODE != kPlain)
  {
    part = block_sum<Acc, true>(part, red);
    if (tid == 0)
      partials[bi] = part;
  }
}

// Kernel variants selectable per matrix: {chunk (entries staged per block),
// vec (consecutive entries per vector load), unroll (vector steps in flight),
// non-temporal matrix loads, threads per workgroup (= rows per block at
// most), short column indices}.  Variants 0-4 also take complex values.
struct SpmvVariant
{
  int chunk;
  int vec;
  int unroll;
  bool nt;
  int bs;
  bool i16;
};
#define SPMVX_VARIANTS(X)                                                                    \
  X(0, 2048, 1, 1, false, 256, false)                                                        \
  X(1, 2048, 1, 4, true, 256, false)                                                         \
  X(2, 2048, 2, 2, true, 256, false)                                                         \
  X(3, 2048, 4, 1, true, 256, false)                                                         \
  X(4, 2048, 4, 2, true, 256, false)                                                         \
  X(5, 8192, 2, 4, true, 1024, false)                                                        \
  X(6, 8192, 2, 2, true, 1024, false)                                                        \
  X(7, 8192, 4, 2, true, 1024, false)                                                        \
  X(8, 8192, 2, 4, true, 1024, true)                                                         \
  X(9, 2048, 4, 1, true, 256, true)                                                          \
  X(10, 4096, 8, 1, true, 512, true)                                                         \
  X(11, 4096, 4, 2, true, 512, true)
constexpr int kNumVariants = 12;
#define SPMVX_VARIANT_ENTRY(i, C, V, U, N, B, I) {C, V, U, N, B, I},
constexpr SpmvVariant kVariants[kNumVariants] = {SPMVX_VARIANTS(SPMVX_VARIANT_ENTRY)};
#undef SPMVX_VARIANT_ENTRY

// Entries and rows per row block of a variant.  The sliced layout
// (kSlicedVariant, defined below) reports the blocks of variant 0, which
// CSR-only operations (transpose, SDDMM) then use.
inline int variant_block_nnz(int var)
{
  if (var < 0 || var >= kNumVariants)
    var = 0;
  return kVariants[var].chunk;
}
inline int variant_block_rows(int var)
{
  if (var < 0 || var >= kNumVariants)
    var = 0;
  return kVariants[var].bs;
}
inline bool variant_short(int var) { return var >= 0 && var < kNumVariants && kVariants[var].i16; }

// Device arrays of a CSR matrix for one row blocking: the CSR arrays, the
// block descriptors and, for the short-index variants, the 16-bit column
// indices and the per-block window bases (null when not built).
struct CsrView
{
  const int32_t* rowptr;
  const int32_t* colidx;
  const int2* blocks;
  int64_t nblocks;
  const uint16_t* col16;
  const int4* wbase;
  int32_t ncols;
  int short_kind; // 1: one base per block, 2: four windows per block
};

// Launches kernel with args on stream s; when ev0 / ev1 are given they
// record the kernel's own start and end (hipExtLaunchKernel), so that their
// elapsed time is the kernel's execution without the dispatch before it.
template <typename... KArgs, typename... Args>
inline void launch_timed(void (*kernel)(KArgs...), dim3 grid, dim3 block, uint32_t shmem,
                         hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, Args... args)
{
  if (ev0 || ev1)
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, ev0, ev1, 0, static_cast<KArgs>(args)...);
  else
    kernel<<<grid, block, shmem, s>>>(static_cast<KArgs>(args)...);
}

// Launches the SpMV kernel of variant `var` (blocks built with
// variant_block_nnz/rows(var)); returns the number of partial sums written
// (MODE != kPlain).  long_par: rows longer than a block are summed in
// parallel (long_row_sum) instead of in stored order.  ev0 / ev1: see
// launch_timed.
template <typename V, typename X, typename Acc, int MODE>
int64_t launch_csr_stream(int var, const CsrView& M, hipStream_t s, const V* val, const X* x,
                          X* y, const X* b, Acc* partials, const CgState<Acc>* st, int32_t it,
                          int32_t long_par = 0, hipEvent_t ev0 = nullptr,
                          hipEvent_t ev1 = nullptr)
{
  if (M.nblocks <= 0)
    return 0;
  const dim3 grid(static_cast<unsigned>(M.nblocks));
#define SPMVX_LAUNCH_IM(IMV)                                                                 \
  launch_timed(&csr_stream<V, X, Acc, MODE, C, VV, U, N, B, IMV>, grid, dim3(B), 0, s, ev0,   \
               ev1, M.rowptr, M.colidx, M.col16, M.wbase, M.ncols, val, x, y, M.blocks,       \
               M.nblocks, b, partials, st, it, long_par)
#define SPMVX_CASE(i, CC, VVV, UU, NN, BB, I)                                                \
  case i:                                                                                    \
  {                                                                                          \
    constexpr int C = CC, VV = VVV, U = UU, B = BB;                                          \
    constexpr bool N = NN;                                                                   \
    if constexpr (kIsComplex<V> && (C > 2048 || I))                                          \
      throw Error(SPMVX_ERR_ARG, "kernel variant not available for complex values");         \
    else if constexpr (I)                                                                    \
    {                                                                                        \
      if (M.short_kind == 1)                                                                 \
        SPMVX_LAUNCH_IM(1);                                                                  \
      else                                                                                   \
        SPMVX_LAUNCH_IM(2);                                                                  \
    }                                                                                        \
    else                                                                                     \
      SPMVX_LAUNCH_IM(0);                                                                    \
    break;                                                                                   \
  }
  switch (var)
  {
    SPMVX_VARIANTS(SPMVX_CASE)
  default: break;
  }
#undef SPMVX_CASE
#undef SPMVX_LAUNCH_IM
  return M.nblocks;
}

// ---------------------------------------------------------------- sliced row patterns
// Layout for matrices whose rows repeat a few column-offset patterns
// (stencils, structured meshes): no column indices are stored.  Each row keeps
// the 16-bit id of its pattern (the offsets col - row of its entries, in the
// row's CSR order) and its values; the values of 64 consecutive rows (a
// slice, one wavefront) are interleaved so that entry k of all 64 rows is one
// coalesced load: element (row lane, entry k) of a slice sits at
//   slice_base + (k / VW) * 64 * VW + lane * VW + k % VW,
// the slice width (longest row) rounded up to VW and padded with zeros.
// Patterns live in LDS as rows of (length, offset_0 .. offset_{wmax-1}).
// Entries are summed in CSR order from 0 with separately rounded products, so
// every y entry is bit-identical to csr_stream's.
constexpr int kSlice = 64;
constexpr int kSlicedVariant = SPMVX_SLICED_VARIANT;  // kernel variant id of this layout
static_assert(kSlicedVariant == kNumVariants, "the sliced layout follows the CSR variants");

// Entries per lane per value load (VW) for values of vbytes bytes: one for
// every value type (no padding: the slice width stays the longest row).
// Measured on the 7- and 27-point stencils: FP32 with 4-byte loads ran
// 12-18% faster than with 8-byte loads of value pairs (7 entries padded to
// 8), BF16 with 2-byte loads 4-6% faster than with 8-byte loads of four.
__host__ __device__ constexpr int sliced_vw_of(int /* vbytes */)
{
  return 1;
}
template <typename V>
constexpr int sliced_vw()
{
  return sliced_vw_of(static_cast<int>(sizeof(V)));
}

// One wavefront handles RPL consecutive slices (64 x RPL rows, lane l owns
// row l of each) so that the value loads and gathers of all of them are in
// flight together; each step issues UNR value loads (VW entries each) per
// slice before any product, and the launcher picks UNR from the slice width
// (sliced_step).
constexpr int kSliceBS = 256;
template <typename V, typename X, typename Acc, int MODE, int UNR, int RPL>
__global__ __launch_bounds__(kSliceBS) void sliced_spmv(
    const uint16_t* __restrict__ pid, const int2* __restrict__ sdesc, int64_t nslices,
    int64_t nrows, int64_t ncols, const int32_t* __restrict__ dict, int32_t dict_n,
    int32_t wmax, const V* __restrict__ val, const X* __restrict__ x, X* __restrict__ y,
    const X* __restrict__ b, Acc* __restrict__ partials, const CgState<Acc>* st, int32_t it)
{
  if (MODE != kPlain && !cg_live(st, it))
    return;
  constexpr int VW = sliced_vw<V>();
  constexpr int SPW = (kSliceBS / kSlice) * RPL; // slices per workgroup
  extern __shared__ int32_t sdict[];             // dict_n entries
  __shared__ Acc red[16];
  const int64_t nblk = (nslices + SPW - 1) / SPW;
  const int64_t bi = xcd_block(blockIdx.x, nblk);
  const int64_t s0 = bi * SPW + (threadIdx.x >> 6) * RPL; // this wavefront's first slice
  const int lane = threadIdx.x & 63;
  // slice descriptors, pattern ids and the row's x (p.y) or b (residual)
  // entry are loaded before the pattern table is staged, so that they are in
  // flight while the workgroup waits for the table
  int2 d[RPL];
  int p[RPL];
  bool has[RPL];
  X xr[RPL];
#pragma unroll
  for (int j = 0; j < RPL; ++j)
  {
    const int64_t sl = s0 + j;
    const bool sv = sl < nslices;
    d[j] = sv ? sdesc[sl] : make_int2(0, 0); // W = 0: loads stay at slot 0
    const int64_t row = sl * kSlice + lane;
    has[j] = sv && row < nrows;
    p[j] = has[j] ? static_cast<int>(pid[row]) : 0;
    xr[j] = X(0);
    if constexpr (MODE == kDot)
      xr[j] = has[j] ? x[row] : X(0);
    else if constexpr (MODE == kResid)
      xr[j] = has[j] ? b[row] : X(0);
  }
  for (int i = threadIdx.x; i < dict_n; i += kSliceBS)
    sdict[i] = dict[i];
  lds_barrier();
  Acc part = Acc(0);
  if (s0 < nslices) // uniform per wavefront
  {
    int W = 0;
    int wj[RPL], len[RPL];
    const int32_t* pat[RPL];
    const V* vs[RPL];
    int32_t r32[RPL], rc[RPL];
    Acc sum[RPL];
#pragma unroll
    for (int j = 0; j < RPL; ++j)
    {
      const int64_t sl = s0 + j;
      wj[j] = d[j].y;
      W = d[j].y > W ? d[j].y : W;
      const int64_t row = sl * kSlice + lane;
      pat[j] = sdict + p[j] * (wmax + 1);
      len[j] = has[j] ? pat[j][0] : 0;
      // column used by idle slots: any valid index (ncols >= 1 when nnz >= 1)
      rc[j] = static_cast<int32_t>(row < ncols ? row : ncols - 1);
      r32[j] = static_cast<int32_t>(row);
      vs[j] = val + static_cast<int64_t>(d[j].x) * kSlice + lane * VW;
      sum[j] = Acc(0);
    }
    for (int k0 = 0; k0 < W; k0 += VW * UNR)
    {
      V v[RPL][UNR][VW];
      X xv[RPL][UNR][VW];
#pragma unroll
      for (int j = 0; j < RPL; ++j)
#pragma unroll
        for (int u = 0; u < UNR; ++u)
        {
          const int kk = k0 + u * VW;
          const int kl = kk < wj[j] ? kk : 0;
          load_vec<true>(vs[j] + (kl / VW) * (kSlice * VW), v[j][u]);
        }
#pragma unroll
      for (int j = 0; j < RPL; ++j)
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
          for (int e = 0; e < VW; ++e)
          {
            const int k = k0 + u * VW + e;
            const bool ok = k < len[j];
            const int32_t o = pat[j][1 + (ok ? k : 0)]; // stays inside the pattern row
            xv[j][u][e] = x[ok ? r32[j] + o : rc[j]];
          }
#pragma unroll
      for (int j = 0; j < RPL; ++j)
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
          for (int e = 0; e < VW; ++e)
          {
            const int k = k0 + u * VW + e;
            const Acc pr = static_cast<Acc>(widen(v[j][u][e])) * static_cast<Acc>(xv[j][u][e]);
            sum[j] = (k < len[j]) ? sum[j] + pr : sum[j];
          }
    }
#pragma unroll
    for (int j = 0; j < RPL; ++j)
      if (has[j])
      {
        const int64_t row = (s0 + j) * kSlice + lane;
        if (MODE == kResid)
        {
          const X yi = xr[j] - static_cast<X>(sum[j]);
          store_result(y + row, yi);
          part += static_cast<Acc>(yi) * static_cast<Acc>(yi);
        }
        else
        {
          const X yi = static_cast<X>(sum[j]);
          store_result(y + row, yi);
          if (MODE == kDot)
            part += static_cast<Acc>(xr[j]) * static_cast<Acc>(yi);
        }
      }
  }
  if (MODE == kPlain)
    return;
  part = block_sum<Acc, true>(part, red);
  if (threadIdx.x == 0)
    partials[bi] = part;
}

// Device arrays of a sliced layout (values excluded).
struct SlicedView
{
  const uint16_t* pid;
  const int2* sdesc;
  int64_t nslices, nrows, ncols;
  const int32_t* dict;
  int32_t dict_n, wmax;
  int rpl; // slices per wavefront: 1, 2 or 4
};

// Launches the sliced kernel; returns the number of partial sums written.
// Slots per step for a slice of w slots: among the candidate step sizes,
// the one that leaves the fewest idle slots in the last step, then the one
// with the fewest steps.
inline int sliced_step(int w, std::initializer_list<int> sizes)
{
  int best = 0, best_idle = 0, best_steps = 0;
  for (int u : sizes)
  {
    const int steps = (w + u - 1) / u;
    const int idle = steps * u - w;
    if (best == 0 || idle < best_idle || (idle == best_idle && steps < best_steps))
    {
      best = u;
      best_idle = idle;
      best_steps = steps;
    }
  }
  return best;
}

// Value loads per step (UNR) of the sliced kernel for values of vbytes bytes,
// rpl slices per wavefront and the padded pattern width wpad (slots): at
// each value type's default rpl, the compiled step (UNR x VW slots) that
// leaves the fewest idle slots, then the fewest steps (the best of the sizes
// compiled on each of the 27- and 7-point stencils and banded matrices of
// widths 5-41 measured; a step of exactly 7 entries took the 7-point
// stencil's FP64 product from 281 to 260 us and its FP32 one from 118 to
// 114 us); with rpl set by hand, the fixed step.
inline int sliced_unr(int vbytes, int rpl, int wpad)
{
  if (vbytes == 8)
  {
    if (rpl == 1)
      return sliced_step(wpad, {4, 7, 8, 10, 14});
    return 4;
  }
  if (vbytes == 4)
  {
    if (rpl == 2)
      return sliced_step(wpad, {4, 7, 8, 14});
    return 4;
  }
  if (rpl == 2)
    return sliced_step(wpad, {4, 7});
  return 4;
}

template <typename V, typename X, typename Acc, int MODE>
int64_t launch_sliced(const SlicedView& S, hipStream_t s, const V* val, const X* x, X* y,
                      const X* b, Acc* partials, const CgState<Acc>* st, int32_t it,
                      hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr)
{
  const size_t lds = sizeof(int32_t) * static_cast<size_t>(S.dict_n);
  const int spw = (kSliceBS / kSlice) * S.rpl;
  const int64_t nblk = (S.nslices + spw - 1) / spw;
  if (nblk <= 0)
    return 0;
  const dim3 grid(static_cast<unsigned>(nblk));
  auto go = [&](auto unr, auto rpl) {
    launch_timed(&sliced_spmv<V, X, Acc, MODE, decltype(unr)::value, decltype(rpl)::value>, grid,
                 dim3(kSliceBS), static_cast<uint32_t>(lds), s, ev0, ev1, S.pid, S.sdesc,
                 S.nslices, S.nrows, S.ncols, S.dict, S.dict_n, S.wmax, val, x, y, b, partials, st,
                 it);
  };
  using std::integral_constant;
  constexpr int VW = sliced_vw<V>();
  const int wpad = (S.wmax + VW - 1) / VW * VW; // slice width in slots
  const int u = sliced_unr(static_cast<int>(sizeof(V)), S.rpl, wpad);
  if constexpr (sizeof(V) == 8)
  {
    if (S.rpl == 1)
    {
      if (u == 14)
        go(integral_constant<int, 14>{}, integral_constant<int, 1>{});
      else if (u == 10)
        go(integral_constant<int, 10>{}, integral_constant<int, 1>{});
      else if (u == 8)
        go(integral_constant<int, 8>{}, integral_constant<int, 1>{});
      else if (u == 7)
        go(integral_constant<int, 7>{}, integral_constant<int, 1>{});
      else
        go(integral_constant<int, 4>{}, integral_constant<int, 1>{});
    }
    else if (S.rpl == 2)
      go(integral_constant<int, 4>{}, integral_constant<int, 2>{});
    else
      go(integral_constant<int, 4>{}, integral_constant<int, 4>{});
  }
  else if constexpr (sizeof(V) == 4)
  {
    if (S.rpl == 2)
    {
      if (u == 14)
        go(integral_constant<int, 14>{}, integral_constant<int, 2>{});
      else if (u == 8)
        go(integral_constant<int, 8>{}, integral_constant<int, 2>{});
      else if (u == 7)
        go(integral_constant<int, 7>{}, integral_constant<int, 2>{});
      else
        go(integral_constant<int, 4>{}, integral_constant<int, 2>{});
    }
    else if (S.rpl == 4)
      go(integral_constant<int, 4>{}, integral_constant<int, 4>{});
    else
      go(integral_constant<int, 4>{}, integral_constant<int, 1>{});
  }
  else
  {
    if (S.rpl == 2)
    {
      if (u == 7)
        go(integral_constant<int, 7>{}, integral_constant<int, 2>{});
      else
        go(integral_constant<int, 4>{}, integral_constant<int, 2>{});
    }
    else if (S.rpl == 4)
      go(integral_constant<int, 4>{}, integral_constant<int, 4>{});
    else
      go(integral_constant<int, 4>{}, integral_constant<int, 1>{});
  }
  return nblk;
}

// ---------------------------------------------------------------- CG vectors
constexpr int kVecBS = 256;

// 16 bytes of a real vector (4 floats or 2 doubles) for the streaming vector
// kernels: with one element per lane and load, a grid-stride pass keeps too
// few bytes in flight per CU to reach HBM bandwidth (measured on 16.8M-row
// CG: the float updates run 6-7% faster with 16-byte passes; with the cache
// policy below, FP64 CG 1.8% faster).
template <typename X>
struct Pack16
{
  static constexpr int N = 16 / static_cast<int>(sizeof(X));
  X v[N];
};

// NT: a non-temporal access (streamed, not kept in the caches)
template <bool NT = false, typename X>
__device__ __forceinline__ Pack16<X> ld16(const X* p)
{
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  Pack16<X> r;
  const u4 q = ld_stream<NT>(reinterpret_cast<const u4*>(p));
  __builtin_memcpy(r.v, &q, 16);
  return r;
}

template <bool NT = false, typename X>
__device__ __forceinline__ void st16(X* p, const Pack16<X>& r)
{
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  u4 q;
  __builtin_memcpy(&q, r.v, 16);
  if constexpr (NT)
    __builtin_nontemporal_store(q, reinterpret_cast<u4*>(p));
  else
    *reinterpret_cast<u4*>(p) = q;
}

// Cache policy of the CG vectors.  Per iteration the product gathers p and
// writes y, the residual update reads y and rewrites r, the last update reads
// r and rewrites p and x: x is touched by that one kernel only, so its
// accesses stream (non-temporal) and leave the last-level cache to p, r and
// y; for doubles the last reads of y (residual update) and of r (x/p update)
// in an iteration stream as well.  Measured on 16.8M-row CG: x streamed took
// FP64 CG 3-4% and FP32 refinement 5% faster; y streamed as well another 3%
// for FP64 and nothing for FP32, r another 1-2% for FP64 and 1-1.5% slower
// for FP32; p's load streamed made both slower.
template <typename X>
constexpr bool kStreamLastRead = sizeof(X) == 8;

__device__ __forceinline__ bool aligned16(const void* a, const void* b, const void* c = nullptr)
{
  return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)
           | reinterpret_cast<uintptr_t>(c)) & 15u) == 0;
}

// x = 0, r = b, p[0:n) = b, p[n:ncols) = 0, partial ||b||^2 (cg.cpp:31-44)
template <typename X, typename Acc>
__global__ __launch_bounds__(kVecBS) void cg_init(int64_t n, int64_t ncols,
                                                  const X* __restrict__ b, X* __restrict__ x,
                                                  X* __restrict__ r, X* __restrict__ p,
                                                  Acc* __restrict__ partials)
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
  if (threadIdx.x == 0)
    partials[blockIdx.x] = part;
}

// rnorm0 = rnorm_old = sum, k = 0 (cg.cpp:42-47)
template <typename Acc>
__global__ __launch_bounds__(kFinBS) void cg_init_finalize(const Acc* __restrict__ partials,
                                                           int64_t n, CgState<Acc>* st,
                                                           Acc* __restrict__ hist)
{
  __shared__ Acc red[16];
  const Acc s = finalize_sum(partials, n, red);
  if (threadIdx.x == 0)
  {
    st->rnorm0 = s;
    st->rnorm_old = s;
    st->k = 0;
    st->converged = 0;
    st->xp_pending = 0;
    st->stopped = st->kmax > 0 ? 0 : 1;
    if (hist)
      hist[0] = s;
  }
}

// alpha = rnorm_old / (p.y)  (cg.cpp:59-62).  Also retires the previous
// iteration's x/p update (which ran before this kernel in stream order).
template <typename Acc>
__global__ __launch_bounds__(kFinBS) void cg_alpha_finalize(const Acc* __restrict__ partials,
                                                            int64_t n, CgState<Acc>* st,
                                                            int32_t it)
{
  const bool live = cg_live(st, it);
  __syncthreads();
  if (threadIdx.x == 0)
    st->xp_pending = 0;
  if (!live)
    return;
  __shared__ Acc red[16];
  const Acc s = finalize_sum(partials, n, red);
  if (threadIdx.x == 0)
    st->alpha = st->rnorm_old / s;
}

// last = sum (refinement residual norm)
template <typename Acc>
__global__ __launch_bounds__(kFinBS) void sum_finalize(const Acc* __restrict__ partials,
                                                       int64_t n, CgState<Acc>* st)
{
  __shared__ Acc red[16];
  const Acc s = finalize_sum(partials, n, red);
  if (threadIdx.x == 0)
    st->last = s;
}

// r -= alpha y; partial ||r||^2 (cg.cpp:66,69)
template <typename X, typename Acc>
__global__ __launch_bounds__(kVecBS) void cg_update_r(int64_t n, X* __restrict__ r,
                                                      const X* __restrict__ y,
                                                      Acc* __restrict__ partials,
                                                      const CgState<Acc>* st, int32_t it)
{
  if (!cg_live(st, it))
    return;
  __shared__ Acc red[16];
  const X alpha = static_cast<X>(st->alpha);
  Acc part = Acc(0);
  constexpr int N = Pack16<X>::N;
  const int64_t gid = blockIdx.x * static_cast<int64_t>(kVecBS) + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBS;
  int64_t done = 0; // elements covered by the 16-byte passes
  if (aligned16(r, y))
  {
    // two 16-byte packs per thread and step, both loaded before either is
    // used (the second re-reads the first when past the end; stored only
    // when in range)
    const int64_t npk = n / N;
    for (int64_t q = gid; q < npk; q += 2 * stride)
    {
      const bool two = q + stride < npk;
      const int64_t q2 = two ? q + stride : q;
      Pack16<X> ra = ld16(r + q * N), ya = ld16<kStreamLastRead<X>>(y + q * N);
      Pack16<X> rb = ld16(r + q2 * N), yb = ld16<kStreamLastRead<X>>(y + q2 * N);
#pragma unroll
      for (int e = 0; e < N; ++e)
      {
        ra.v[e] = ra.v[e] - alpha * ya.v[e];
        part += static_cast<Acc>(ra.v[e]) * static_cast<Acc>(ra.v[e]);
      }
      st16(r + q * N, ra);
      if (two)
      {
#pragma unroll
        for (int e = 0; e < N; ++e)
        {
          rb.v[e] = rb.v[e] - alpha * yb.v[e];
          part += static_cast<Acc>(rb.v[e]) * static_cast<Acc>(rb.v[e]);
        }
        st16(r + q2 * N, rb);
      }
    }
    done = npk * N;
  }
  for (int64_t i = done + gid; i < n; i += stride)
  {
    const X ri = r[i] - alpha * ld_stream<kStreamLastRead<X>>(y + i);
    r[i] = ri;
    part += static_cast<Acc>(ri) * static_cast<Acc>(ri);
  }
  part = block_sum(part, red);
  if (threadIdx.x == 0)
    partials[blockIdx.x] = part;
}

// rnorm_new, beta, k = it, convergence (cg.cpp:69-79)
template <typename Acc>
__global__ __launch_bounds__(kFinBS) void cg_beta_finalize(const Acc* __restrict__ partials,
                                                           int64_t n, CgState<Acc>* st,
                                                           Acc* __restrict__ hist, int32_t it)
{
  if (!cg_live(st, it))
    return;
  __shared__ Acc red[16];
  const Acc s = finalize_sum(partials, n, red);
  if (threadIdx.x == 0)
  {
    const int32_t k = st->k + 1;
    st->beta = s / st->rnorm_old;
    st->rnorm_old = s;
    st->k = k;
    if (hist)
      hist[k] = s;
    st->converged = (s / st->rnorm0 < st->rtol2) ? 1 : 0;
    st->stopped = (st->converged || k >= st->kmax) ? 1 : 0;
    st->xp_pending = 1;
  }
}

// x += alpha p; p = beta p + r (cg.cpp:65,76): runs iff this iteration
// completed its residual update (also on the iteration that stops).
template <typename X, typename Acc>
__global__ __launch_bounds__(kVecBS) void cg_update_xp(int64_t n, X* __restrict__ x,
                                                       X* __restrict__ p,
                                                       const X* __restrict__ r,
                                                       const CgState<Acc>* st, int32_t it)
{
  (void)it;
  if (!st->xp_pending)
    return;
  const X alpha = static_cast<X>(st->alpha);
  const X beta = static_cast<X>(st->beta);
  constexpr int N = Pack16<X>::N;
  const int64_t gid = blockIdx.x * static_cast<int64_t>(kVecBS) + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBS;
  int64_t done = 0; // elements covered by the 16-byte passes
  if (aligned16(x, p, r))
  {
    const int64_t npk = n / N;
    for (int64_t q = gid; q < npk; q += stride)
    {
      Pack16<X> xa = ld16<true>(x + q * N), pa = ld16(p + q * N);
      const Pack16<X> ra = ld16<kStreamLastRead<X>>(r + q * N);
#pragma unroll
      for (int e = 0; e < N; ++e)
      {
        const X pi = pa.v[e];
        xa.v[e] = xa.v[e] + alpha * pi;
        pa.v[e] = pi * beta + ra.v[e];
      }
      st16<true>(x + q * N, xa);
      st16(p + q * N, pa);
    }
    done = npk * N;
  }
  for (int64_t i = done + gid; i < n; i += stride)
  {
    const X pi = p[i];
    __builtin_nontemporal_store(ld_stream<true>(x + i) + alpha * pi, x + i);
    p[i] = pi * beta + ld_stream<kStreamLastRead<X>>(r + i);
  }
}

} // namespace spmvx
