// Transposed copy of a matrix for Matrix<T>::transpmult (spmv/Matrix.cpp:
// 206-211), built on the device.  Eigen evaluates A^T x as a scatter over the
// rows of A in ascending order, so every output entry y_c accumulates the
// products of column c in ascending row order; the transposed CSR keeps
// exactly that order (rows ascending within each of its rows), and the
// ordered CSR kernels then sum them in it (bit-identical).
//
//   1. stable radix sort of the entries by column (keys: column indices,
//      values: entry positions, which ascend with the row): perm[d] is the
//      position in A of the d-th entry of A^T;
//   2. transposed row pointer from the sorted columns (run boundaries);
//   3. transposed column indices: the row of entry perm[d] (binary search in
//      A's row pointer); values: gathered through perm.
// perm is kept, so a value update re-gathers the values only.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "handles.hpp"

namespace spmvx
{
namespace
{

// Row of every entry: the last r with rowptr[r] <= k (rows may be empty).
__global__ void entry_rows(const int32_t* __restrict__ rowptr, int64_t nrows, int64_t nnz,
                           const int32_t* __restrict__ perm, int32_t* __restrict__ tcol)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; d < nnz; d += stride)
  {
    const int32_t k = perm[d];
    int64_t lo = 0, hi = nrows; // rowptr[lo] <= k < rowptr[hi]
    while (hi - lo > 1)
    {
      const int64_t mid = (lo + hi) >> 1;
      if (rowptr[mid] <= k)
        lo = mid;
      else
        hi = mid;
    }
    tcol[d] = static_cast<int32_t>(lo);
  }
}

// trowptr[c] = number of entries with column < c, from the sorted columns:
// position d starts every column in (col[d-1], col[d]] (d = nnz: up to ncols).
__global__ void run_starts(const uint32_t* __restrict__ scol, int64_t nnz, int64_t ncols,
                           int32_t* __restrict__ trowptr)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; d <= nnz; d += stride)
  {
    const int64_t lo = d == 0 ? 0 : static_cast<int64_t>(scol[d - 1]) + 1;
    const int64_t hi = d == nnz ? ncols : static_cast<int64_t>(scol[d]);
    for (int64_t c = lo; c <= hi; ++c)
      trowptr[c] = static_cast<int32_t>(d);
  }
}

template <typename E>
__global__ void gather_entries(const int32_t* __restrict__ perm, const E* __restrict__ src,
                               int64_t nnz, E* __restrict__ dst)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t d = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; d < nnz; d += stride)
    dst[d] = src[perm[d]];
}

int grid_of(int64_t n)
{
  int64_t g = (n + 255) / 256;
  return static_cast<int>(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

struct alignas(16) u128
{
  uint64_t lo, hi;
};

} // namespace

void gather_transposed_values(spmvx_matrix_s* A, hipStream_t s)
{
  const int64_t nnz = A->pat->nnz;
  if (nnz == 0)
    return;
  const size_t es = dtype_size(A->dtype);
  const int32_t* perm = A->tperm.as<int32_t>();
  const int g = grid_of(nnz);
  if (es == 16)
    gather_entries<<<g, 256, 0, s>>>(perm, A->values.as<u128>(), nnz, A->tvalues.as<u128>());
  else if (es == 8)
    gather_entries<<<g, 256, 0, s>>>(perm, A->values.as<uint64_t>(), nnz, A->tvalues.as<uint64_t>());
  else if (es == 4)
    gather_entries<<<g, 256, 0, s>>>(perm, A->values.as<uint32_t>(), nnz, A->tvalues.as<uint32_t>());
  else
    gather_entries<<<g, 256, 0, s>>>(perm, A->values.as<uint16_t>(), nnz, A->tvalues.as<uint16_t>());
  check_launch("gather_entries");
}

void build_transpose(spmvx_matrix_s* A, hipStream_t s)
{
  const Pattern& P = *A->pat;
  const int64_t nnz = P.nnz, nrows = P.nrows, ncols = P.ncols;
  auto T = std::make_shared<Pattern>();
  T->nrows = ncols;
  T->ncols = nrows;
  T->nnz = nnz;
  T->device = P.device;
  T->rowptr.ensure(sizeof(int32_t) * static_cast<size_t>(ncols + 1));
  T->colidx.ensure(sizeof(int32_t) * static_cast<size_t>(nnz + 16)); // zero padding
  A->tperm.ensure(sizeof(int32_t) * static_cast<size_t>(nnz > 0 ? nnz : 1));
  A->tvalues.ensure(dtype_size(A->dtype) * static_cast<size_t>(nnz + 16));
  DevBuf scol, tmp; // freed after the stream has drained
  scol.ensure(sizeof(uint32_t) * static_cast<size_t>(nnz > 0 ? nnz : 1));
  if (nnz > 0)
  {
    // columns < ncols < 2^31: sort on the bits they use
    unsigned bits = 1;
    while (bits < 31 && (int64_t(1) << bits) < ncols)
      ++bits;
    const auto* keys = reinterpret_cast<const uint32_t*>(P.colidx.ptr);
    rocprim::counting_iterator<int32_t> pos(0);
    size_t tmp_bytes = 0;
    SPMVX_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, scol.as<uint32_t>(), pos,
                                        A->tperm.as<int32_t>(), static_cast<size_t>(nnz), 0, bits,
                                        s));
    tmp.ensure(tmp_bytes > 0 ? tmp_bytes : 1);
    SPMVX_HIP(rocprim::radix_sort_pairs(tmp.ptr, tmp_bytes, keys, scol.as<uint32_t>(), pos,
                                        A->tperm.as<int32_t>(), static_cast<size_t>(nnz), 0, bits,
                                        s));
    entry_rows<<<grid_of(nnz), 256, 0, s>>>(P.rowptr.as<int32_t>(), nrows, nnz,
                                            A->tperm.as<int32_t>(), T->colidx.as<int32_t>());
    check_launch("entry_rows");
    gather_transposed_values(A, s);
    run_starts<<<grid_of(nnz + 1), 256, 0, s>>>(scol.as<uint32_t>(), nnz, ncols,
                                                T->rowptr.as<int32_t>());
    check_launch("run_starts");
  }
  // the host keeps the row pointer (row-binning analysis of A^T); without
  // entries it stays as allocated, zero
  T->h_rowptr.assign(static_cast<size_t>(ncols + 1), 0);
  if (nnz > 0)
    SPMVX_HIP(hipMemcpyAsync(T->h_rowptr.data(), T->rowptr.ptr,
                             sizeof(int32_t) * static_cast<size_t>(ncols + 1),
                             hipMemcpyDeviceToHost, s));
  SPMVX_HIP(hipStreamSynchronize(s));
  (void)variant_blocks(*T, default_variant(A->dtype));
  A->tpat = std::move(T);
}

} // namespace spmvx
