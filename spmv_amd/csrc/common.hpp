// Shared host/device helpers of libspmv_amd: error state, HIP checks,
// reductions and the XCD-aware block mapping.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "spmv_amd.h"

namespace spmvx
{

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error
{
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define SPMVX_HIP(call)                                                         \
  do                                                                            \
  {                                                                             \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      throw ::spmvx::Error(SPMVX_ERR_HIP, std::string(#call) + ": "             \
                                              + hipGetErrorString(e_));         \
  } while (0)

#define SPMVX_REQUIRE(cond, msg)                                                \
  do                                                                            \
  {                                                                             \
    if (!(cond))                                                                \
      throw ::spmvx::Error(SPMVX_ERR_ARG, msg);                                 \
  } while (0)

// Runs `body` and converts exceptions into a status code + last error.
template <class F>
int guarded(F&& body)
{
  try
  {
    set_last_error("");
    body();
    return SPMVX_OK;
  }
  catch (const Error& e)
  {
    set_last_error(e.what());
    return e.code;
  }
  catch (const std::bad_alloc&)
  {
    set_last_error("out of host memory");
    return SPMVX_ERR_ARG;
  }
  catch (const std::exception& e)
  {
    set_last_error(e.what());
    return SPMVX_ERR_ARG;
  }
}

inline hipStream_t as_stream(spmvx_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Checks that the last kernel launch was accepted.
inline void check_launch(const char* what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    throw Error(SPMVX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- device
// gfx950: 8 XCDs, workgroups dispatched round-robin (workgroup i -> XCD i%8).
constexpr int kNumXcd = 8;

// Bijection of dispatch order onto logical block index such that each XCD
// processes one contiguous range of logical blocks (so neighbouring rows,
// which share x entries, meet in the same XCD's L2).
__device__ __forceinline__ int64_t xcd_block(int64_t i, int64_t n)
{
  const int64_t x = i % kNumXcd;
  const int64_t j = i / kNumXcd;
  const int64_t q = n / kNumXcd, rem = n % kNumXcd;
  const int64_t start = x * q + (x < rem ? x : rem);
  return start + j;
}

// Wavefront (64 lanes) sum via cross-lane shuffles; fixed order.
template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off, 64);
  return v;
}

// Workgroup barrier ordering LDS accesses only: global loads and stores
// still in flight stay in flight (__syncthreads would wait for all of them,
// which stalls a kernel that has prefetched its next data or just stored
// its results).
__device__ __forceinline__ void lds_barrier()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Block sum (blockDim.x multiple of 64, <= 1024); result valid in all
// threads.  `scratch` (LDS) holds >= 16 elements.  Deterministic order.
// LdsOnly: the barriers order LDS only (see lds_barrier).
template <typename T, bool LdsOnly = false>
__device__ __forceinline__ T block_sum(T v, T* scratch)
{
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nw = blockDim.x >> 6;
  v = wave_sum(v);
  if (LdsOnly)
    lds_barrier();
  else
    __syncthreads();
  if (lane == 0)
    scratch[wid] = v;
  if (LdsOnly)
    lds_barrier();
  else
    __syncthreads();
  T s = (lane < nw) ? scratch[lane] : T(0);
  s = wave_sum(s);
  return s;
}

} // namespace spmvx
