// Matrix handle: CSR upload + row-binning analysis, SpMV, transpose SpMV,
// SDDMM, value updates (spmv/Matrix.h:30-53, spmv/Matrix.cpp:13-211).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <string>
#include <unordered_map>

#include "handles.hpp"

namespace spmvx
{

std::vector<int2> row_blocks(const int32_t* rowptr, int64_t nrows, int64_t max_nnz,
                             int64_t max_rows)
{
  std::vector<int2> b;
  b.reserve(static_cast<size_t>(nrows / 16 + 2));
  b.push_back(make_int2(0, rowptr[0]));
  int64_t r = 0;
  while (r < nrows)
  {
    const int64_t start = r;
    int64_t acc = 0;
    while (r < nrows && r - start < max_rows)
    {
      const int64_t len = rowptr[r + 1] - rowptr[r];
      if (acc + len > max_nnz)
        break;
      acc += len;
      ++r;
    }
    if (r == start)
      ++r; // a single row longer than max_nnz
    b.push_back(make_int2(static_cast<int32_t>(r), rowptr[r]));
  }
  return b;
}

const RowBlocks& Pattern::blocks_for(int chunk, int rows)
{
  std::lock_guard<std::recursive_mutex> lock(lazy_mu);
  const std::pair<int, int> key(chunk, rows);
  auto it = blocks.find(key);
  if (it != blocks.end())
    return *it->second;
  auto rb = std::make_unique<RowBlocks>();
  std::vector<int2> b = row_blocks(h_rowptr.data(), nrows, chunk, rows);
  rb->nblocks = static_cast<int64_t>(b.size()) - 1;
  rb->desc.ensure(sizeof(int2) * b.size());
  SPMVX_HIP(hipMemcpy(rb->desc.ptr, b.data(), sizeof(int2) * b.size(), hipMemcpyHostToDevice));
  const RowBlocks& ref = *rb;
  blocks[key] = std::move(rb);
  return ref;
}

const RowBlocks* Pattern::find_blocks(int chunk, int rows)
{
  std::lock_guard<std::recursive_mutex> lock(lazy_mu);
  auto it = blocks.find(std::pair<int, int>(chunk, rows));
  return it == blocks.end() ? nullptr : it->second.get();
}

const RowPatterns* Pattern::row_patterns()
{
  std::lock_guard<std::recursive_mutex> lock(lazy_mu);
  if (rpat_tried)
    return rpat.get();
  rpat_tried = true;
  if (nnz == 0 || nrows == 0)
  {
    rpat_error = "the sliced layout needs a matrix with entries";
    return nullptr;
  }
  std::vector<int32_t> ci(static_cast<size_t>(nnz));
  SPMVX_HIP(hipMemcpy(ci.data(), colidx.ptr, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
  auto R = std::make_unique<RowPatterns>();
  R->h_pid.resize(static_cast<size_t>(nrows));
  // pattern table under construction: offsets of pattern q at start[q]
  std::vector<int32_t> offs, start;
  std::unordered_map<uint64_t, std::vector<int32_t>> buckets;
  int32_t wmax = 0;
  for (int64_t r = 0; r < nrows; ++r)
  {
    const int32_t a = h_rowptr[r], e = h_rowptr[r + 1];
    const int32_t len = e - a;
    if (len > kMaxSliceWidth)
    {
      rpat_error = "a row is longer than the sliced layout allows (" + std::to_string(len) + " > "
                   + std::to_string(kMaxSliceWidth) + ")";
      return nullptr;
    }
    uint64_t h = 1469598103934665603ull ^ static_cast<uint64_t>(len);
    for (int32_t k = a; k < e; ++k)
      h = (h ^ static_cast<uint64_t>(static_cast<uint32_t>(ci[k] - static_cast<int32_t>(r))))
          * 1099511628211ull;
    std::vector<int32_t>& cand = buckets[h];
    int32_t id = -1;
    for (int32_t q : cand)
      if (R->h_len[q] == len
          && std::equal(offs.begin() + start[q], offs.begin() + start[q] + len, ci.begin() + a,
                        [r](int32_t o, int32_t c) { return o == c - static_cast<int32_t>(r); }))
      {
        id = q;
        break;
      }
    if (id < 0)
    {
      id = static_cast<int32_t>(R->h_len.size());
      if (id >= kMaxPatterns)
      {
        rpat_error = "more than " + std::to_string(kMaxPatterns)
                     + " distinct row patterns (the sliced layout suits stencil-like matrices)";
        return nullptr;
      }
      cand.push_back(id);
      start.push_back(static_cast<int32_t>(offs.size()));
      R->h_len.push_back(len);
      for (int32_t k = a; k < e; ++k)
        offs.push_back(ci[k] - static_cast<int32_t>(r));
      wmax = std::max(wmax, len);
      if (static_cast<int64_t>(R->h_len.size()) * (wmax + 1) > kMaxDictEntries)
      {
        rpat_error = "the row-pattern table exceeds " + std::to_string(kMaxDictEntries)
                     + " entries (too many or too long patterns for the sliced layout)";
        return nullptr;
      }
    }
    R->h_pid[r] = static_cast<uint16_t>(id);
  }
  R->npat = static_cast<int32_t>(R->h_len.size());
  R->wmax = std::max(wmax, 1);
  R->dict_n = R->npat * (R->wmax + 1);
  std::vector<int32_t> dict(static_cast<size_t>(R->dict_n), 0);
  for (int32_t q = 0; q < R->npat; ++q)
  {
    dict[q * (R->wmax + 1)] = R->h_len[q];
    for (int32_t k = 0; k < R->h_len[q]; ++k)
      dict[q * (R->wmax + 1) + 1 + k] = offs[start[q] + k];
  }
  R->dict.ensure(sizeof(int32_t) * dict.size());
  SPMVX_HIP(hipMemcpy(R->dict.ptr, dict.data(), sizeof(int32_t) * dict.size(), hipMemcpyHostToDevice));
  R->pid.ensure(sizeof(uint16_t) * static_cast<size_t>(nrows));
  SPMVX_HIP(hipMemcpy(R->pid.ptr, R->h_pid.data(), sizeof(uint16_t) * nrows, hipMemcpyHostToDevice));
  rpat = std::move(R);
  return rpat.get();
}

const SliceLayout& Pattern::slice_layout(int vw)
{
  std::lock_guard<std::recursive_mutex> lock(lazy_mu);
  auto it = slices.find(vw);
  if (it != slices.end())
    return *it->second;
  const RowPatterns* R = row_patterns();
  SPMVX_REQUIRE(R != nullptr, rpat_error);
  auto L = std::make_unique<SliceLayout>();
  L->nslices = (nrows + kSlice - 1) / kSlice;
  std::vector<int2> d(static_cast<size_t>(L->nslices));
  int64_t off64 = 0;
  for (int64_t sl = 0; sl < L->nslices; ++sl)
  {
    int32_t w = 0;
    const int64_t r1 = std::min<int64_t>(nrows, (sl + 1) * kSlice);
    for (int64_t r = sl * kSlice; r < r1; ++r)
      w = std::max(w, R->h_len[R->h_pid[r]]);
    w = (w + vw - 1) / vw * vw;
    SPMVX_REQUIRE(off64 < (int64_t(1) << 31), "sliced layout too large for 32-bit slice offsets");
    d[sl] = make_int2(static_cast<int32_t>(off64), w);
    off64 += w;
  }
  L->slots = off64 * kSlice;
  L->sdesc.ensure(sizeof(int2) * d.size());
  SPMVX_HIP(hipMemcpy(L->sdesc.ptr, d.data(), sizeof(int2) * d.size(), hipMemcpyHostToDevice));
  const SliceLayout& ref = *L;
  slices[vw] = std::move(L);
  return ref;
}

namespace
{

// Short-index analysis, one thread per row block of at most `chunk` entries
// (longer blocks hold one row each, which the kernels sum from the 32-bit
// indices): the block's column span, and windows of 16384 columns opened
// greedily over its entries in stored order (a new window starts 4096
// columns below the first column none of the open ones holds).  wb1 gets
// the block's first column, wb4 up to four window bases (unused ones repeat
// the first); flags bit 0: some block spans 65536 columns or more, bit 1:
// some block needs more than four windows.
constexpr int32_t kWindow = 16384;
__global__ void short_analyse(const int32_t* __restrict__ colidx, const int2* __restrict__ blocks,
                              int64_t nblocks, int32_t chunk, int32_t* __restrict__ wb1,
                              int4* __restrict__ wb4, int* flags)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t bi = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; bi < nblocks;
       bi += stride)
  {
    const int32_t k0 = blocks[bi].y, k1 = blocks[bi + 1].y;
    int32_t lo = 0x7fffffff, hi = -1;
    int32_t b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    int nw = 0;
    const bool small = k1 - k0 <= chunk;
    for (int32_t k = k0; small && k < k1; ++k)
    {
      const int32_t c = colidx[k];
      lo = c < lo ? c : lo;
      hi = c > hi ? c : hi;
      const bool in = (nw > 0 && c >= b0 && c - b0 < kWindow) || (nw > 1 && c >= b1 && c - b1 < kWindow)
                      || (nw > 2 && c >= b2 && c - b2 < kWindow)
                      || (nw > 3 && c >= b3 && c - b3 < kWindow);
      if (in)
        continue;
      const int32_t nb = c > 4096 ? c - 4096 : 0;
      if (nw == 0)
        b0 = nb;
      else if (nw == 1)
        b1 = nb;
      else if (nw == 2)
        b2 = nb;
      else if (nw == 3)
        b3 = nb;
      ++nw;
    }
    if (nw < 2)
      b1 = b0;
    if (nw < 3)
      b2 = b0;
    if (nw < 4)
      b3 = b0;
    wb1[bi] = hi >= 0 ? lo : 0;
    wb4[bi] = make_int4(b0, b1, b2, b3);
    if (small && hi >= 0 && hi - lo >= 65536)
      atomicOr(flags, 1);
    if (small && nw > 4)
      atomicOr(flags, 2);
  }
}

// Encodes the entries of every block of at most `chunk` entries (one
// workgroup per block): kind 1, the offset from the block's first column;
// kind 2, (window << 14) | offset, the first window holding the column.
// Kind 1 leaves the block's base in wbase.x.
__global__ void short_encode(const int32_t* __restrict__ colidx, const int2* __restrict__ blocks,
                             int64_t nblocks, int32_t chunk, int kind,
                             const int32_t* __restrict__ wb1, int4* __restrict__ wbase,
                             uint16_t* __restrict__ col16)
{
  for (int64_t bi = blockIdx.x; bi < nblocks; bi += gridDim.x)
  {
    const int32_t k0 = blocks[bi].y, k1 = blocks[bi + 1].y;
    if (k1 - k0 > chunk)
      continue;
    const int4 w = wbase[bi];
    const int32_t base = wb1[bi];
    __syncthreads(); // every thread has read wbase[bi] before thread 0 rewrites it
    for (int32_t k = k0 + static_cast<int32_t>(threadIdx.x); k < k1; k += blockDim.x)
    {
      const int32_t c = colidx[k];
      if (kind == 1)
      {
        col16[k] = static_cast<uint16_t>(c - base);
        continue;
      }
      uint32_t sel = 3, off = static_cast<uint32_t>(c - w.w);
      if (c >= w.z && c - w.z < kWindow)
      {
        sel = 2;
        off = static_cast<uint32_t>(c - w.z);
      }
      if (c >= w.y && c - w.y < kWindow)
      {
        sel = 1;
        off = static_cast<uint32_t>(c - w.y);
      }
      if (c >= w.x && c - w.x < kWindow)
      {
        sel = 0;
        off = static_cast<uint32_t>(c - w.x);
      }
      col16[k] = static_cast<uint16_t>((sel << 14) | off);
    }
    if (kind == 1 && threadIdx.x == 0)
      wbase[bi] = make_int4(base, base, base, base);
  }
}

} // namespace

void build_short_index(Pattern& P, const RowBlocks& rbc, int chunk)
{
  std::lock_guard<std::recursive_mutex> lock(P.lazy_mu);
  RowBlocks& rb = const_cast<RowBlocks&>(rbc); // owned by P, guarded by its lock
  if (rb.short_built)
    return;
  rb.short_kind = 0;
  if (rb.nblocks > 0 && P.nnz > 0)
  {
    DevBuf wb1, flags;
    wb1.ensure(sizeof(int32_t) * static_cast<size_t>(rb.nblocks));
    flags.ensure(sizeof(int));
    rb.wbase.ensure(sizeof(int4) * static_cast<size_t>(rb.nblocks));
    const int g = static_cast<int>(std::min<int64_t>((rb.nblocks + 255) / 256, 4096));
    short_analyse<<<g, 256>>>(P.colidx.as<int32_t>(), rb.desc.as<int2>(), rb.nblocks, chunk,
                              wb1.as<int32_t>(), rb.wbase.as<int4>(), flags.as<int>());
    check_launch("short_analyse");
    int f = 0;
    SPMVX_HIP(hipMemcpy(&f, flags.ptr, sizeof(int), hipMemcpyDeviceToHost));
    const int kind = !(f & 1) ? 1 : (!(f & 2) ? 2 : 0);
    if (kind)
    {
      rb.col16.ensure(sizeof(uint16_t) * static_cast<size_t>(P.nnz + 16));
      const int ge = static_cast<int>(std::min<int64_t>(rb.nblocks, 65536));
      short_encode<<<ge, 256>>>(P.colidx.as<int32_t>(), rb.desc.as<int2>(), rb.nblocks, chunk,
                                kind, wb1.as<int32_t>(), rb.wbase.as<int4>(),
                                rb.col16.as<uint16_t>());
      check_launch("short_encode");
      SPMVX_HIP(hipDeviceSynchronize());
    }
    else
      rb.wbase.release();
    rb.short_kind = kind;
  }
  rb.short_built = true;
}

CsrView csr_view(Pattern& P, int var)
{
  const RowBlocks& rb = variant_blocks(P, var);
  const bool sh = variant_short(var);
  if (sh)
  {
    build_short_index(P, rb, variant_block_nnz(var));
    SPMVX_REQUIRE(rb.short_kind != 0 || P.nnz == 0,
                  "the matrix does not fit 16-bit column indices (a row block's columns "
                  "fall into more than four windows of 16384)");
  }
  return CsrView{P.rowptr.as<int32_t>(), P.colidx.as<int32_t>(), rb.desc.as<int2>(), rb.nblocks,
                 sh ? rb.col16.as<uint16_t>() : nullptr, sh ? rb.wbase.as<int4>() : nullptr,
                 static_cast<int32_t>(P.ncols), sh ? rb.short_kind : 0};
}

CsrView csr_view_built(Pattern& P, int var)
{
  const RowBlocks* rb = built_variant_blocks(P, var);
  const bool sh = variant_short(var);
  if (!rb || (sh && !rb->short_built))
    return CsrView{P.rowptr.as<int32_t>(), P.colidx.as<int32_t>(), nullptr, 0, nullptr, nullptr,
                   static_cast<int32_t>(P.ncols), 0};
  return CsrView{P.rowptr.as<int32_t>(), P.colidx.as<int32_t>(), rb->desc.as<int2>(), rb->nblocks,
                 sh ? rb->col16.as<uint16_t>() : nullptr, sh ? rb->wbase.as<int4>() : nullptr,
                 static_cast<int32_t>(P.ncols), sh ? rb->short_kind : 0};
}

// Whether variant `var` of P can read 16-bit column indices (builds the
// analysis).
bool fits_short(Pattern& P, int var)
{
  const RowBlocks& rb = variant_blocks(P, var);
  build_short_index(P, rb, variant_block_nnz(var));
  return rb.short_kind != 0;
}

// Variant override from the environment (SPMVX_SPMV_VARIANT), -1 if none.
int forced_variant()
{
  static int forced = [] {
    const char* e = std::getenv("SPMVX_SPMV_VARIANT");
    if (!e)
      return -1;
    const int x = std::atoi(e);
    return (x >= 0 && x < kNumVariants) ? x : -1;
  }();
  return forced;
}

int default_variant(int32_t dtype)
{
  // measured best 32-bit-index CSR kernels on gfx950 for the stencil
  // matrices: FP64 1024-thread workgroups staging 8192 entries (vec 2 x
  // unroll 4), FP32 vec 4, BF16 vec 4 x unroll 2; SPMVX_SPMV_VARIANT overrides
  const int forced = forced_variant();
  if (dtype == SPMVX_C64 || dtype == SPMVX_C32) // variants 0-4 only (16 KB+ LDS products)
    return forced >= 0 && forced <= 4 ? forced : 2;
  if (forced >= 0)
    return forced;
  return dtype == SPMVX_F64 ? 5 : (dtype == SPMVX_F32 ? 3 : 4);
}

int matrix_default_variant(spmvx_matrix_s* A)
{
  const int base = default_variant(A->dtype);
  if (forced_variant() >= 0 || is_complex(A->dtype))
    return base;
  // the short-index kernel for the dtype when the matrix fits it (the
  // analysis of the inspector, Matrix.cpp:43): FP64 the 16-bit counterpart
  // of its default; FP32 and BF16 512-thread workgroups staging 4096 entries,
  // FP32 with vec 4 x unroll 2 (variant 11), BF16 with 16-byte value loads
  // (vec 8, variant 10): 1-5% (FP32) and 4-7% (BF16) faster per product on
  // the 7- and 27-point stencils than 2048-entry blocks of 256 threads
  const int sv = base == 5 ? 8 : (base == 3 ? 11 : (base == 4 ? 10 : -1));
  Pattern& P = *A->pat;
  if (sv < 0 || P.nnz == 0)
    return base;
  return fits_short(P, sv) ? sv : base;
}

namespace
{

// double -> bf16, round to nearest even, via round-to-odd to float (exact
// double rounding: float keeps >= 2 more bits than bf16).
uint16_t host_f64_to_bf16(double d)
{
  if (std::isnan(d))
    return 0x7FC0;
  float f = static_cast<float>(d);
  if (std::isfinite(f) && static_cast<double>(f) != d)
  {
    if (std::fabs(static_cast<double>(f)) > std::fabs(d))
      f = std::nextafter(f, 0.0f);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u |= 1u;
    std::memcpy(&f, &u, 4);
  }
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

__device__ __forceinline__ uint16_t dev_f64_to_bf16(double d)
{
  if (d != d)
    return 0x7FC0;
  float f = __double2float_rz(d);
  uint32_t u = __float_as_uint(f);
  if (static_cast<double>(f) != d && (u & 0x7F800000u) != 0x7F800000u)
    u |= 1u;
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}



__global__ void convert_values(const double* __restrict__ src, void* __restrict__ dst,
                               int64_t n, int32_t dtype)
{
  // n values; complex dtypes read and write (re, im) pairs
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int64_t m = (dtype == SPMVX_C64 || dtype == SPMVX_C32) ? 2 * n : n;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < m; i += stride)
  {
    const double d = src[i];
    if (dtype == SPMVX_F64 || dtype == SPMVX_C64)
      static_cast<double*>(dst)[i] = d;
    else if (dtype == SPMVX_F32 || dtype == SPMVX_C32)
      static_cast<float*>(dst)[i] = static_cast<float>(d);
    else
      static_cast<bf16_t*>(dst)[i].bits = dev_f64_to_bf16(d);
  }
}

// Copy of n doubles: non-temporal 16-byte loads and stores, four per thread
// in flight per step (the read+write ceiling, and the known-size kernel that
// calibrates the HBM byte counters).
constexpr int kCopyUnroll = 4;
__global__ __launch_bounds__(256) void copy_f64_kernel(double* __restrict__ dst,
                                                       const double* __restrict__ src, int64_t n)
{
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2* s2 = reinterpret_cast<const d2*>(src);
  d2* o2 = reinterpret_cast<d2*>(dst);
  const int64_t n2 = n / 2;
  const int64_t tile = static_cast<int64_t>(gridDim.x) * 256 * kCopyUnroll;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * 256 * kCopyUnroll; base < n2; base += tile)
  {
    d2 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u)
    {
      const int64_t i = base + u * 256 + threadIdx.x;
      v[u] = i < n2 ? __builtin_nontemporal_load(s2 + i) : d2{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u)
    {
      const int64_t i = base + u * 256 + threadIdx.x;
      if (i < n2)
        __builtin_nontemporal_store(v[u], o2 + i);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1))
    dst[n - 1] = src[n - 1];
}

// Streaming read of n doubles: non-temporal 16-byte loads, four per thread in
// flight per step, the values summed (one atomic add per workgroup into *sum)
// so that no load is dead code.  The read ceiling of the SpMV kernels, which
// read ~97% of their bytes.
constexpr int kReadUnroll = 4;
__global__ __launch_bounds__(256) void read_f64_kernel(const double* __restrict__ src, int64_t n,
                                                       double* __restrict__ sum)
{
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2* s2 = reinterpret_cast<const d2*>(src);
  const int64_t n2 = n / 2;
  const int64_t tile = static_cast<int64_t>(gridDim.x) * 256 * kReadUnroll;
  d2 acc = {0.0, 0.0};
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * 256 * kReadUnroll; base < n2; base += tile)
  {
    d2 v[kReadUnroll];
#pragma unroll
    for (int u = 0; u < kReadUnroll; ++u)
    {
      const int64_t i = base + u * 256 + threadIdx.x;
      v[u] = i < n2 ? __builtin_nontemporal_load(s2 + i) : d2{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < kReadUnroll; ++u)
      acc += v[u];
  }
  double t = acc.x + acc.y;
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1))
    t += src[n - 1];
  __shared__ double red[16];
  t = block_sum(t, red);
  if (threadIdx.x == 0)
    atomicAdd(sum, t);
}

// dst[slot(row, k)] = src[rowptr[row] + k]: CSR values into the sliced
// row-pattern layout (padding slots keep the zeros they were allocated with).
template <typename T>
__global__ void scatter_sliced(const int32_t* __restrict__ rowptr,
                               const uint16_t* __restrict__ pid,
                               const int32_t* __restrict__ dict, int32_t wmax,
                               const int2* __restrict__ sdesc, int64_t nrows, int vw,
                               const T* __restrict__ src, T* __restrict__ dst)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; r < nrows;
       r += stride)
  {
    const int64_t sl = r / kSlice;
    const int lane = static_cast<int>(r % kSlice);
    const int64_t base = static_cast<int64_t>(sdesc[sl].x) * kSlice;
    const int len = dict[static_cast<int64_t>(pid[r]) * (wmax + 1)];
    const int32_t a = rowptr[r];
    for (int k = 0; k < len; ++k)
      dst[base + (k / vw) * (kSlice * vw) + lane * vw + k % vw] = src[a + k];
  }
}

// out[k] = alpha * u[row(k)] * v[col(k)]; one workgroup per row block, rows of
// an entry found by binary search in the block's row pointers (LDS).
__global__ __launch_bounds__(kBS) void sddmm_kernel(const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ colidx,
                                                    const int2* __restrict__ blocks,
                                                    int64_t nblocks, double alpha,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ out)
{
  __shared__ int32_t rp[kBS + 1];
  const int64_t bi = xcd_block(blockIdx.x, nblocks);
  const int32_t r0 = blocks[bi].x, r1 = blocks[bi + 1].x;
  const int nr = r1 - r0;
  for (int t = threadIdx.x; t <= nr; t += kBS)
    rp[t] = rowptr[r0 + t];
  lds_barrier();
  const int32_t k0 = rp[0], k1 = rp[nr];
  for (int32_t k = k0 + threadIdx.x; k < k1; k += kBS)
  {
    int lo = 0, hi = nr - 1; // last row with rp[row] <= k
    while (lo < hi)
    {
      const int mid = (lo + hi + 1) >> 1;
      if (rp[mid] <= k)
        lo = mid;
      else
        hi = mid - 1;
    }
    out[k] = (alpha * u[r0 + lo]) * v[colidx[k]];
  }
}

template <typename V, typename X, typename Acc>
void launch_plain(int var, Pattern& P, const void* vals, const void* x, void* y, hipStream_t s,
                  int32_t long_par, hipEvent_t ev0, hipEvent_t ev1)
{
  const CsrView M = csr_view(P, var);
  launch_csr_stream<V, X, Acc, kPlain>(var, M, s, static_cast<const V*>(vals),
                                       static_cast<const X*>(x), static_cast<X*>(y), nullptr,
                                       nullptr, nullptr, 0, long_par, ev0, ev1);
  check_launch("csr SpMV");
}

// Bytes of one vector element of a matrix of `dtype` (BF16 matrices take
// FP32 vectors).
size_t vec_bytes(int32_t dtype)
{
  return dtype == SPMVX_BF16 ? 4 : dtype_size(dtype);
}

// w[0:n) = 0 (a kernel, so a caller may capture it in a graph)
__global__ void zero_words(uint32_t* __restrict__ w, int64_t n)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    w[i] = 0u;
}

// ev0 / ev1 (optional): events that record the product kernel's own start
// and end (recorded around the zero fill when there are no entries).
void launch_spmv(int32_t dtype, int var, Pattern& P, const void* vals, const void* x, void* y,
                 hipStream_t s, int32_t long_par, hipEvent_t ev0 = nullptr,
                 hipEvent_t ev1 = nullptr)
{
  if (P.nnz == 0)
  {
    // no stored entries: y = 0.  The product kernels gather x at padding
    // columns without a branch, which needs x to hold at least one entry;
    // a matrix without entries may have no columns at all.
    const int64_t words = P.nrows * static_cast<int64_t>(vec_bytes(dtype)) / 4;
    if (ev0)
      SPMVX_HIP(hipEventRecord(ev0, s));
    if (words > 0)
    {
      zero_words<<<vec_grid(words), kVecBS, 0, s>>>(static_cast<uint32_t*>(y), words);
      check_launch("zero_words");
    }
    if (ev1)
      SPMVX_HIP(hipEventRecord(ev1, s));
    return;
  }
  if (dtype == SPMVX_F64)
    launch_plain<double, double, double>(var, P, vals, x, y, s, long_par, ev0, ev1);
  else if (dtype == SPMVX_F32)
    launch_plain<float, float, float>(var, P, vals, x, y, s, long_par, ev0, ev1);
  else if (dtype == SPMVX_BF16)
    launch_plain<bf16_t, float, float>(var, P, vals, x, y, s, long_par, ev0, ev1);
  else if (dtype == SPMVX_C64)
    launch_plain<cplx<double>, cplx<double>, cplx<double>>(var, P, vals, x, y, s, 0, ev0, ev1);
  else
    launch_plain<cplx<float>, cplx<float>, cplx<float>>(var, P, vals, x, y, s, 0, ev0, ev1);
}

// Fills A->svalues from A->values on stream s (layout built beforehand).
void scatter_values(spmvx_matrix_s* A, const RowPatterns& R, const SliceLayout& L, hipStream_t s)
{
  const Pattern& P = *A->pat;
  const size_t es = dtype_size(A->dtype);
  const int vw = sliced_vw_of(static_cast<int>(es));
  const int64_t g64 = std::min<int64_t>((P.nrows + 255) / 256, 8192);
  const dim3 grid(static_cast<unsigned>(g64 > 0 ? g64 : 1)), block(256);
#define SPMVX_SCATTER(T)                                                                    \
  scatter_sliced<T><<<grid, block, 0, s>>>(P.rowptr.as<int32_t>(), R.pid.as<uint16_t>(),    \
                                           R.dict.as<int32_t>(), R.wmax, L.sdesc.as<int2>(), \
                                           P.nrows, vw, A->values.as<T>(), A->svalues.as<T>())
  if (es == 8)
    SPMVX_SCATTER(uint64_t);
  else if (es == 4)
    SPMVX_SCATTER(uint32_t);
  else
    SPMVX_SCATTER(uint16_t);
#undef SPMVX_SCATTER
  check_launch("scatter_sliced");
}

std::shared_ptr<Pattern> upload_pattern(int64_t nrows, int64_t ncols, int64_t nnz,
                                        const int32_t* rowptr, const int32_t* colidx)
{
  auto P = std::make_shared<Pattern>();
  P->nrows = nrows;
  P->ncols = ncols;
  P->nnz = nnz;
  SPMVX_HIP(hipGetDevice(&P->device));
  P->rowptr.ensure(sizeof(int32_t) * static_cast<size_t>(nrows + 1));
  SPMVX_HIP(hipMemcpy(P->rowptr.ptr, rowptr, sizeof(int32_t) * (nrows + 1), hipMemcpyHostToDevice));
  P->colidx.ensure(sizeof(int32_t) * static_cast<size_t>(nnz + 16));
  if (nnz > 0)
    SPMVX_HIP(hipMemcpy(P->colidx.ptr, colidx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  P->h_rowptr.assign(rowptr, rowptr + nrows + 1);
  (void)variant_blocks(*P, default_variant(SPMVX_F64));
  return P;
}

void validate_csr(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* rowptr,
                  const int32_t* colidx)
{
  SPMVX_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0, "negative matrix size");
  // entry offsets in the kernels run up to one staging step (at most 64 Ki
  // entries) past nnz in 32-bit arithmetic
  SPMVX_REQUIRE(nrows < (int64_t(1) << 31) && ncols < (int64_t(1) << 31)
                    && nnz <= (int64_t(1) << 31) - (int64_t(1) << 16),
                "matrix too large for 32-bit indices");
  SPMVX_REQUIRE(rowptr != nullptr, "rowptr is NULL");
  SPMVX_REQUIRE(nnz == 0 || colidx != nullptr, "colidx is NULL");
  SPMVX_REQUIRE(rowptr[0] == 0, "rowptr[0] must be 0");
  SPMVX_REQUIRE(rowptr[nrows] == nnz, "rowptr[nrows] must equal nnz");
  for (int64_t i = 0; i < nrows; ++i)
    SPMVX_REQUIRE(rowptr[i + 1] >= rowptr[i], "rowptr must be non-decreasing");
  for (int64_t k = 0; k < nnz; ++k)
    SPMVX_REQUIRE(colidx[k] >= 0 && colidx[k] < ncols, "column index out of range");
}

} // namespace

void wait_values(spmvx_matrix_s* A)
{
  if (A->values_pending)
  {
    SPMVX_HIP(hipEventSynchronize(A->values_event));
    A->values_pending = false;
  }
}

SlicedView sliced_view(spmvx_matrix_s* A)
{
  SPMVX_REQUIRE(!is_complex(A->dtype), "the sliced layout is for real values");
  Pattern& P = *A->pat;
  const int vw = sliced_vw_of(static_cast<int>(dtype_size(A->dtype)));
  const RowPatterns* R = P.row_patterns();
  SPMVX_REQUIRE(R != nullptr, P.rpat_error);
  const SliceLayout& L = P.slice_layout(vw);
  if (!A->svalues_ready)
  {
    wait_values(A); // the scatter below runs on the null stream
    A->svalues.release();
    A->svalues.ensure(dtype_size(A->dtype) * static_cast<size_t>(L.slots + 16));
    scatter_values(A, *R, L, nullptr);
    SPMVX_HIP(hipDeviceSynchronize());
    A->svalues_ready = true;
  }
  // slices per wavefront: SPMVX_SLICED_RPL (1, 2 or 4) overrides the default
  static const int forced = [] {
    const char* e = std::getenv("SPMVX_SLICED_RPL");
    const int v = e ? std::atoi(e) : 0;
    return (v == 1 || v == 2 || v == 4) ? v : 0;
  }();
  // measured on the 27- and 7-point stencils: FP64 one slice per wavefront,
  // FP32 and BF16 two
  const int bs = forced ? forced : (A->dtype == SPMVX_F64 ? 1 : 2);
  return SlicedView{R->pid.as<uint16_t>(), L.sdesc.as<int2>(), L.nslices, P.nrows, P.ncols,
                    R->dict.as<int32_t>(),  R->dict_n,          R->wmax,   bs};
}

namespace
{

// CSR kernel variant of the transposed product: the matrix's, unless it is
// the sliced layout or a 16-bit-index kernel that the transposed pattern
// does not fit (then the dtype's 32-bit default).
int transposed_variant(spmvx_matrix_s* A)
{
  const int v = A->variant;
  if (v == kSlicedVariant || (variant_short(v) && !fits_short(*A->tpat, v)))
    return default_variant(A->dtype);
  return v;
}

// ev0 / ev1 (optional): events that record the product kernel's own start
// and end.
void launch_matrix_spmv(spmvx_matrix_s* A, const void* x, void* y, hipStream_t s,
                        hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr)
{
  if (A->variant != kSlicedVariant)
  {
    launch_spmv(A->dtype, A->variant, *A->pat, A->values.ptr, x, y, s, A->long_rows, ev0, ev1);
    return;
  }
  const SlicedView v = sliced_view(A);
  if (A->dtype == SPMVX_F64)
    launch_sliced<double, double, double, kPlain>(v, s, A->svalues.as<double>(),
                                                  static_cast<const double*>(x),
                                                  static_cast<double*>(y), nullptr, nullptr,
                                                  nullptr, 0, ev0, ev1);
  else if (A->dtype == SPMVX_F32)
    launch_sliced<float, float, float, kPlain>(v, s, A->svalues.as<float>(),
                                               static_cast<const float*>(x),
                                               static_cast<float*>(y), nullptr, nullptr, nullptr,
                                               0, ev0, ev1);
  else
    launch_sliced<bf16_t, float, float, kPlain>(v, s, A->svalues.as<bf16_t>(),
                                                static_cast<const float*>(x),
                                                static_cast<float*>(y), nullptr, nullptr, nullptr,
                                                0, ev0, ev1);
  check_launch("sliced SpMV");
}

} // namespace
} // namespace spmvx

using namespace spmvx;

extern "C" {

int spmvx_matrix_create(int32_t dtype, int64_t nrows, int64_t ncols, int64_t nnz,
                        const int32_t* rowptr, const int32_t* colidx, const double* values,
                        spmvx_matrix* out)
{
  return guarded([&] {
    SPMVX_REQUIRE(out != nullptr, "out is NULL");
    *out = nullptr;
    const size_t es = dtype_size(dtype);
    validate_csr(nrows, ncols, nnz, rowptr, colidx);
    SPMVX_REQUIRE(nnz == 0 || values != nullptr, "values is NULL");
    auto A = std::make_unique<spmvx_matrix_s>();
    A->dtype = dtype;
    A->pat = upload_pattern(nrows, ncols, nnz, rowptr, colidx);
    A->variant = matrix_default_variant(A.get());
    A->values.ensure(es * static_cast<size_t>(nnz + 16));
    if (nnz > 0)
    {
      if (dtype == SPMVX_F64 || dtype == SPMVX_C64)
        SPMVX_HIP(hipMemcpy(A->values.ptr, values, es * nnz, hipMemcpyHostToDevice));
      else if (dtype == SPMVX_F32 || dtype == SPMVX_C32)
      {
        std::vector<float> f(values, values + (es / 4) * nnz);
        SPMVX_HIP(hipMemcpy(A->values.ptr, f.data(), es * nnz, hipMemcpyHostToDevice));
      }
      else
      {
        std::vector<uint16_t> h(static_cast<size_t>(nnz));
        for (int64_t k = 0; k < nnz; ++k)
          h[k] = host_f64_to_bf16(values[k]);
        SPMVX_HIP(hipMemcpy(A->values.ptr, h.data(), 2 * nnz, hipMemcpyHostToDevice));
      }
    }
    *out = A.release();
  });
}

int spmvx_matrix_create_like(spmvx_matrix base, int32_t dtype, spmvx_matrix* out)
{
  return guarded([&] {
    SPMVX_REQUIRE(base != nullptr && out != nullptr, "NULL handle");
    SPMVX_REQUIRE(base->dtype == SPMVX_F64, "create_like needs an FP64 base matrix");
    *out = nullptr;
    const size_t es = dtype_size(dtype);
    auto A = std::make_unique<spmvx_matrix_s>();
    A->dtype = dtype;
    A->long_rows = base->long_rows;
    A->pat = base->pat;
    A->variant = matrix_default_variant(A.get());
    const int64_t nnz = base->pat->nnz;
    A->values.ensure(es * static_cast<size_t>(nnz + 16));
    wait_values(base); // the conversion below runs on the null stream
    if (nnz > 0)
    {
      convert_values<<<1024, 256>>>(base->values.as<double>(), A->values.ptr, nnz, dtype);
      check_launch("convert_values");
      SPMVX_HIP(hipDeviceSynchronize());
    }
    *out = A.release();
  });
}

int spmvx_matrix_destroy(spmvx_matrix A)
{
  return guarded([&] { delete A; });
}

int spmvx_matrix_info(spmvx_matrix A, int32_t* dtype, int64_t* nrows, int64_t* ncols,
                      int64_t* nnz, int64_t* nblocks)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    if (dtype) *dtype = A->dtype;
    if (nrows) *nrows = A->pat->nrows;
    if (ncols) *ncols = A->pat->ncols;
    if (nnz) *nnz = A->pat->nnz;
    if (nblocks) *nblocks = variant_blocks(*A->pat, A->variant).nblocks;
  });
}

int spmvx_matrix_arrays(spmvx_matrix A, const int32_t** rowptr, const int32_t** colidx,
                        const void** values)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    if (rowptr) *rowptr = A->pat->rowptr.as<int32_t>();
    if (colidx) *colidx = A->pat->colidx.as<int32_t>();
    if (values) *values = A->values.ptr;
  });
}

int spmvx_matrix_set_values(spmvx_matrix A, const double* values_dev, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    const int64_t nnz = A->pat->nnz;
    SPMVX_REQUIRE(nnz == 0 || values_dev != nullptr, "values is NULL");
    if (nnz > 0)
    {
      convert_values<<<1024, 256, 0, as_stream(stream)>>>(values_dev, A->values.ptr, nnz, A->dtype);
      check_launch("convert_values");
      if (A->svalues_ready)
      {
        const int vw = sliced_vw_of(static_cast<int>(dtype_size(A->dtype)));
        scatter_values(A, *A->pat->row_patterns(), A->pat->slice_layout(vw), as_stream(stream));
      }
      if (A->tpat) // the transposed copy keeps its pattern: values re-gathered
        gather_transposed_values(A, as_stream(stream));
      if (!A->values_event)
        SPMVX_HIP(hipEventCreateWithFlags(&A->values_event, hipEventDisableTiming));
      SPMVX_HIP(hipEventRecord(A->values_event, as_stream(stream)));
      A->values_pending = true;
    }
  });
}

int spmvx_matrix_set_kernel(spmvx_matrix A, int32_t variant)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(variant >= -1 && variant <= kSlicedVariant, "unknown kernel variant");
    SPMVX_REQUIRE(!is_complex(A->dtype) || variant <= 4,
                  "kernel variant not available for complex values (use 0-4)");
    const int v = variant < 0 ? matrix_default_variant(A) : variant;
    if (v == kSlicedVariant)
      (void)sliced_view(A); // builds the layout or reports why it does not apply
    else
      (void)csr_view(*A->pat, v); // row blocks (and short indices) built here
    A->variant = v;
  });
}

int spmvx_num_kernels(void) { return kSlicedVariant + 1; }

int spmvx_matrix_set_long_rows(spmvx_matrix A, int32_t parallel)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(parallel == 0 || parallel == 1, "long-row mode must be 0 or 1");
    SPMVX_REQUIRE(parallel == 0 || !is_complex(A->dtype),
                  "parallel long rows are for real values");
    A->long_rows = parallel;
  });
}

int spmvx_matrix_get_kernel(spmvx_matrix A, int32_t* variant, char* name, int32_t name_len)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    const int v = A->variant;
    if (variant)
      *variant = v;
    if (name && name_len > 0)
    {
      std::string d;
      if (v == kSlicedVariant)
      {
        // the slots each step loads (value loads x entries per load), as the
        // launcher picks them from the padded pattern width
        const SlicedView sv = sliced_view(A);
        const int vb = static_cast<int>(dtype_size(A->dtype));
        const int vw = sliced_vw_of(vb);
        const int wpad = (sv.wmax + vw - 1) / vw * vw;
        d = "sliced_spmv (row-pattern slices of 64 rows, " + std::to_string(sv.rpl)
            + " per wavefront, " + std::to_string(sliced_unr(vb, sv.rpl, wpad) * vw)
            + " slots per step)";
      }
      else
      {
        const SpmvVariant& k = kVariants[v];
        d = std::string("csr_stream (") + std::to_string(k.bs) + " threads, chunk "
            + std::to_string(k.chunk) + ", vec " + std::to_string(k.vec) + ", unroll "
            + std::to_string(k.unroll) + (k.nt ? ", non-temporal" : "")
            + (k.i16 ? std::string(", 16-bit column indices, ")
                           + (built_variant_blocks(*A->pat, v)
                                      && built_variant_blocks(*A->pat, v)->short_kind == 1
                                  ? "one base per block"
                                  : "four windows per block")
                     : std::string())
            + ")";
      }
      const size_t n = std::min(d.size(), static_cast<size_t>(name_len - 1));
      std::memcpy(name, d.data(), n);
      name[n] = '\0';
    }
  });
}

int spmvx_matrix_sliced_info(spmvx_matrix A, int32_t* npatterns, int32_t* wmax, int64_t* slots)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(!is_complex(A->dtype), "the sliced layout is for real values");
    Pattern& P = *A->pat;
    const RowPatterns* R = P.row_patterns();
    SPMVX_REQUIRE(R != nullptr, P.rpat_error);
    const SliceLayout& L = P.slice_layout(sliced_vw_of(static_cast<int>(dtype_size(A->dtype))));
    if (npatterns) *npatterns = R->npat;
    if (wmax) *wmax = R->wmax;
    if (slots) *slots = L.slots;
  });
}

namespace
{
// The product reads x while it writes y: the two ranges must not overlap.
void require_disjoint(const void* x, int64_t nx, const void* y, int64_t ny, size_t es)
{
  const auto xa = reinterpret_cast<uintptr_t>(x), ya = reinterpret_cast<uintptr_t>(y);
  const uintptr_t xe = xa + static_cast<uintptr_t>(nx) * es, ye = ya + static_cast<uintptr_t>(ny) * es;
  SPMVX_REQUIRE(nx == 0 || ny == 0 || xe <= ya || ye <= xa, "x and y must not overlap");
}
} // namespace

int spmvx_spmv(spmvx_matrix A, const void* x, void* y, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(A->pat->nrows == 0 || y != nullptr, "y is NULL");
    SPMVX_REQUIRE(A->pat->nnz == 0 || x != nullptr, "x is NULL");
    require_disjoint(x, A->pat->ncols, y, A->pat->nrows, vec_bytes(A->dtype));
    launch_matrix_spmv(A, x, y, as_stream(stream));
  });
}

int spmvx_spmv_timed(spmvx_matrix A, const void* x, void* y, spmvx_stream stream,
                     void* start_event, void* stop_event)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(A->pat->nrows == 0 || y != nullptr, "y is NULL");
    SPMVX_REQUIRE(A->pat->nnz == 0 || x != nullptr, "x is NULL");
    SPMVX_REQUIRE(start_event != nullptr && stop_event != nullptr, "event is NULL");
    require_disjoint(x, A->pat->ncols, y, A->pat->nrows, vec_bytes(A->dtype));
    launch_matrix_spmv(A, x, y, as_stream(stream), static_cast<hipEvent_t>(start_event),
                       static_cast<hipEvent_t>(stop_event));
  });
}

int spmvx_spmv_t(spmvx_matrix A, const void* x, void* y, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(A->pat->ncols == 0 || y != nullptr, "y is NULL");
    SPMVX_REQUIRE(A->pat->nnz == 0 || x != nullptr, "x is NULL");
    require_disjoint(x, A->pat->nrows, y, A->pat->ncols, vec_bytes(A->dtype));
    if (!A->tpat)
      build_transpose(A, as_stream(stream));
    launch_spmv(A->dtype, transposed_variant(A), *A->tpat, A->tvalues.ptr, x, y, as_stream(stream),
                A->long_rows);
  });
}

int spmvx_sddmm(spmvx_matrix A, double alpha, const double* u, const double* v, double* out,
                spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(A->dtype == SPMVX_F64, "sddmm needs an FP64 matrix");
    Pattern& P = *A->pat;
    const RowBlocks& rb = P.blocks_for(2048, kBS); // the kernel stages <= kBS row pointers
    if (rb.nblocks == 0 || P.nnz == 0)
      return;
    SPMVX_REQUIRE(u && v && out, "NULL vector");
    sddmm_kernel<<<dim3(static_cast<unsigned>(rb.nblocks)), dim3(kBS), 0, as_stream(stream)>>>(
        P.rowptr.as<int32_t>(), P.colidx.as<int32_t>(), rb.desc.as<int2>(), rb.nblocks, alpha,
        u, v, out);
    check_launch("sddmm_kernel");
  });
}

int spmvx_copy_f64(double* dst, const double* src, int64_t n, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(n >= 0, "negative size");
    if (n == 0)
      return;
    SPMVX_REQUIRE(dst && src, "NULL pointer");
    SPMVX_REQUIRE((reinterpret_cast<uintptr_t>(dst) & 15) == 0
                      && (reinterpret_cast<uintptr_t>(src) & 15) == 0,
                  "copy needs 16-byte aligned pointers");
    int64_t grid = (n / 2 + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll);
    if (grid > 1024)
      grid = 1024;
    if (grid < 1)
      grid = 1;
    copy_f64_kernel<<<dim3(static_cast<unsigned>(grid)), dim3(256), 0, as_stream(stream)>>>(dst, src, n);
    check_launch("copy_f64_kernel");
  });
}

int spmvx_read_f64(const double* src, int64_t n, double* sum, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(n >= 0, "negative size");
    SPMVX_REQUIRE(sum, "NULL pointer");
    SPMVX_REQUIRE(n == 0 || src, "NULL pointer");
    SPMVX_REQUIRE((reinterpret_cast<uintptr_t>(src) & 15) == 0
                      && (reinterpret_cast<uintptr_t>(sum) & 7) == 0,
                  "read needs a 16-byte aligned source and an 8-byte aligned sum");
    hipStream_t s = as_stream(stream);
    zero_words<<<1, kVecBS, 0, s>>>(reinterpret_cast<uint32_t*>(sum), 2);
    check_launch("zero_words");
    if (n == 0)
      return;
    int64_t grid = (n / 2 + 256 * kReadUnroll - 1) / (256 * kReadUnroll);
    if (grid > 1024)
      grid = 1024;
    if (grid < 1)
      grid = 1;
    read_f64_kernel<<<dim3(static_cast<unsigned>(grid)), dim3(256), 0, s>>>(src, n, sum);
    check_launch("read_f64_kernel");
  });
}

} // extern "C"
