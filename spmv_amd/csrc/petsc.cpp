// Reader of PETSc binary matrix and vector files (spmv/read_petsc.cpp:13-260)
// for one rank of a row-block decomposition.  Host parsing only; the device
// objects are made through the public C ABI (spmvx_matrix_create,
// spmvx_l2gmap_create), so this file has no device code.
//
// File layout (big-endian): matrix = int32 classid 1211216, rows, cols, nnz,
// then int32 row lengths[rows], int32 column indices[nnz], double
// values[nnz]; vector = int32 classid 1211214, rows, then double values[rows].
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "spmv_amd.h"

namespace spmvx
{
void set_last_error(const std::string& msg);
}

struct spmvx_petsc_csr_s
{
  int32_t nranks = 1, rank = 0;
  int64_t nrows_global = 0, ncols_global = 0;
  std::vector<int64_t> row_ranges, col_ranges; // owner_ranges of rows / columns
  std::vector<int32_t> rowptr, colidx;         // local rows, local column numbering
  std::vector<double> values;
  std::vector<int64_t> ghosts;                 // global indices, ascending
  std::vector<int64_t> remote_ptr, remote_ghosts; // the other ranks' ghosts owned here
};

namespace
{

struct ReadError : std::runtime_error
{
  using std::runtime_error::runtime_error;
};

int guarded_host(const std::function<void()>& body)
{
  try
  {
    spmvx::set_last_error("");
    body();
    return SPMVX_OK;
  }
  catch (const ReadError& e)
  {
    spmvx::set_last_error(e.what());
    return SPMVX_ERR_ARG;
  }
  catch (const std::bad_alloc&)
  {
    spmvx::set_last_error("out of host memory");
    return SPMVX_ERR_ARG;
  }
}

// Rows (or columns) 0..N-1 split into `size` near-equal contiguous ranges,
// the first N % size ranges one longer (read_petsc.cpp:13-30).
std::vector<int64_t> owner_ranges(int size, int64_t N)
{
  const int64_t n = N / size, r = N % size;
  std::vector<int64_t> ranges(static_cast<size_t>(size) + 1);
  for (int q = 0; q <= size; ++q)
    ranges[q] = q < r ? q * (n + 1) : q * n + r;
  return ranges;
}

class BigEndianFile
{
public:
  explicit BigEndianFile(const char* path) : f_(path ? std::fopen(path, "rb") : nullptr)
  {
    if (!f_)
      throw ReadError("Could not open file");
    std::fseek(f_, 0, SEEK_END);
    size_ = std::ftell(f_);
    std::fseek(f_, 0, SEEK_SET);
  }
  ~BigEndianFile()
  {
    if (f_)
      std::fclose(f_);
  }
  BigEndianFile(const BigEndianFile&) = delete;
  BigEndianFile& operator=(const BigEndianFile&) = delete;
  int64_t size() const { return size_; }
  void seek(int64_t pos) { std::fseek(f_, static_cast<long>(pos), SEEK_SET); }
  void read_i32(int32_t* out, int64_t n, const char* what)
  {
    raw(out, 4, n, what);
    for (int64_t i = 0; i < n; ++i)
      out[i] = static_cast<int32_t>(__builtin_bswap32(static_cast<uint32_t>(out[i])));
  }
  void read_f64(double* out, int64_t n, const char* what)
  {
    raw(out, 8, n, what);
    for (int64_t i = 0; i < n; ++i)
    {
      uint64_t u;
      std::memcpy(&u, &out[i], 8);
      u = __builtin_bswap64(u);
      std::memcpy(&out[i], &u, 8);
    }
  }

private:
  void raw(void* out, size_t es, int64_t n, const char* what)
  {
    if (n > 0 && std::fread(out, es, static_cast<size_t>(n), f_) != static_cast<size_t>(n))
      throw ReadError(std::string(what) + " is truncated");
  }
  std::FILE* f_;
  int64_t size_ = 0;
};

void parse_matrix(const char* path, int32_t nranks, int32_t rank, spmvx_petsc_csr_s& M)
{
  if (nranks < 1 || rank < 0 || rank >= nranks)
    throw ReadError("invalid rank / nranks");
  BigEndianFile f(path);
  int32_t head[4] = {0, 0, 0, 0};
  if (f.size() < 16)
    throw ReadError("PETSc Matrix file is truncated");
  f.read_i32(head, 4, "PETSc Matrix file");
  if (head[0] != 1211216) // read_petsc.cpp:64-66
    throw ReadError("Bad signature in PETSc Matrix file");
  const int64_t nrows = head[1], ncols = head[2], nnz_tot = head[3];
  if (nrows < 0 || ncols < 0 || nnz_tot < 0)
    throw ReadError("PETSc Matrix file has a negative size");
  if (f.size() < 16 + 4 * nrows + 12 * nnz_tot)
    throw ReadError("PETSc Matrix file is truncated");
  std::vector<int32_t> len(static_cast<size_t>(nrows));
  f.read_i32(len.data(), nrows, "PETSc Matrix file");
  std::vector<int64_t> start(static_cast<size_t>(nrows) + 1, 0);
  for (int64_t i = 0; i < nrows; ++i)
  {
    if (len[i] < 0)
      throw ReadError("PETSc Matrix file has a negative row length");
    start[i + 1] = start[i] + len[i];
  }
  if (start[nrows] != nnz_tot) // read_petsc.cpp:93
    throw ReadError("PETSc Matrix file: row lengths do not add up to the entry count");
  // every column index (the other ranks' ghosts are needed for the plan)
  std::vector<int32_t> cols(static_cast<size_t>(nnz_tot));
  f.read_i32(cols.data(), nnz_tot, "PETSc Matrix file");
  for (int32_t c : cols)
    if (c < 0 || c >= ncols)
      throw ReadError("PETSc Matrix file: column index out of range");

  M.nranks = nranks;
  M.rank = rank;
  M.nrows_global = nrows;
  M.ncols_global = ncols;
  M.row_ranges = owner_ranges(nranks, nrows);
  M.col_ranges = owner_ranges(nranks, ncols);
  const int64_t rr0 = M.row_ranges[rank], rr1 = M.row_ranges[rank + 1];
  const int64_t c0 = M.col_ranges[rank], c1 = M.col_ranges[rank + 1];
  const int64_t ncols_local = c1 - c0;
  const int64_t k0 = start[rr0], k1 = start[rr1];

  // ghosts: off-range columns of the local rows in ascending global order
  // (read_petsc.cpp:115-133)
  M.ghosts.clear();
  for (int64_t k = k0; k < k1; ++k)
    if (cols[k] < c0 || cols[k] >= c1)
      M.ghosts.push_back(cols[k]);
  std::sort(M.ghosts.begin(), M.ghosts.end());
  M.ghosts.erase(std::unique(M.ghosts.begin(), M.ghosts.end()), M.ghosts.end());
  if (ncols_local + static_cast<int64_t>(M.ghosts.size()) >= (int64_t(1) << 31))
    throw ReadError("local column count too large for 32-bit indices");

  // values of the local rows (read_petsc.cpp:105-106,137-141)
  std::vector<double> vals(static_cast<size_t>(k1 - k0));
  f.seek(16 + 4 * nrows + 4 * nnz_tot + 8 * k0);
  f.read_f64(vals.data(), k1 - k0, "PETSc Matrix file");

  // local CSR; every row in ascending local column (Eigen's insert keeps a
  // row sorted, read_petsc.cpp:152-160): owned columns first, then ghosts
  const int64_t nloc = rr1 - rr0;
  M.rowptr.assign(static_cast<size_t>(nloc) + 1, 0);
  M.colidx.resize(static_cast<size_t>(k1 - k0));
  M.values.resize(static_cast<size_t>(k1 - k0));
  std::vector<std::pair<int32_t, double>> row;
  for (int64_t i = 0; i < nloc; ++i)
  {
    const int64_t a = start[rr0 + i], e = start[rr0 + i + 1];
    row.clear();
    for (int64_t k = a; k < e; ++k)
    {
      const int64_t g = cols[k];
      int64_t lc;
      if (g >= c0 && g < c1)
        lc = g - c0;
      else
        lc = ncols_local + (std::lower_bound(M.ghosts.begin(), M.ghosts.end(), g) - M.ghosts.begin());
      row.emplace_back(static_cast<int32_t>(lc), vals[k - k0]);
    }
    std::stable_sort(row.begin(), row.end(),
                     [](const auto& p, const auto& q) { return p.first < q.first; });
    for (size_t j = 1; j < row.size(); ++j)
      if (row[j].first == row[j - 1].first)
        throw ReadError("PETSc Matrix file: duplicate column in row " + std::to_string(rr0 + i));
    for (size_t j = 0; j < row.size(); ++j)
    {
      M.colidx[a - k0 + j] = row[j].first;
      M.values[a - k0 + j] = row[j].second;
    }
    M.rowptr[i + 1] = static_cast<int32_t>(e - k0);
  }

  // the ghosts every other rank takes from this one (what the Alltoall and
  // neighbourhood exchange of L2GMap.cpp:60,112 deliver)
  M.remote_ptr.assign(static_cast<size_t>(nranks) + 1, 0);
  M.remote_ghosts.clear();
  std::vector<int64_t> mine;
  for (int q = 0; q < nranks; ++q)
  {
    if (q != rank)
    {
      mine.clear();
      for (int64_t k = start[M.row_ranges[q]]; k < start[M.row_ranges[q + 1]]; ++k)
        if (cols[k] >= c0 && cols[k] < c1)
          mine.push_back(cols[k]);
      std::sort(mine.begin(), mine.end());
      mine.erase(std::unique(mine.begin(), mine.end()), mine.end());
      M.remote_ghosts.insert(M.remote_ghosts.end(), mine.begin(), mine.end());
    }
    M.remote_ptr[q + 1] = static_cast<int64_t>(M.remote_ghosts.size());
  }
}

} // namespace

extern "C" {

int spmvx_petsc_csr_read(const char* path, int32_t nranks, int32_t rank, spmvx_petsc_csr* out)
{
  return guarded_host([&] {
    if (!out)
      throw ReadError("out is NULL");
    *out = nullptr;
    auto M = std::make_unique<spmvx_petsc_csr_s>();
    parse_matrix(path, nranks, rank, *M);
    *out = M.release();
  });
}

int spmvx_petsc_csr_destroy(spmvx_petsc_csr h)
{
  return guarded_host([&] { delete h; });
}

int spmvx_petsc_csr_info(spmvx_petsc_csr h, int64_t* nrows_global, int64_t* ncols_global,
                         int64_t* nrows_local, int64_t* ncols_local, int64_t* nghosts,
                         int64_t* nnz_local)
{
  return guarded_host([&] {
    if (!h)
      throw ReadError("NULL handle");
    if (nrows_global) *nrows_global = h->nrows_global;
    if (ncols_global) *ncols_global = h->ncols_global;
    if (nrows_local) *nrows_local = h->row_ranges[h->rank + 1] - h->row_ranges[h->rank];
    if (ncols_local) *ncols_local = h->col_ranges[h->rank + 1] - h->col_ranges[h->rank];
    if (nghosts) *nghosts = static_cast<int64_t>(h->ghosts.size());
    if (nnz_local) *nnz_local = static_cast<int64_t>(h->colidx.size());
  });
}

int spmvx_petsc_csr_arrays(spmvx_petsc_csr h, const int32_t** rowptr, const int32_t** colidx,
                           const double** values, const int64_t** ghosts)
{
  return guarded_host([&] {
    if (!h)
      throw ReadError("NULL handle");
    if (rowptr) *rowptr = h->rowptr.data();
    if (colidx) *colidx = h->colidx.data();
    if (values) *values = h->values.data();
    if (ghosts) *ghosts = h->ghosts.data();
  });
}

int spmvx_petsc_csr_remote(spmvx_petsc_csr h, const int64_t** remote_ptr,
                           const int64_t** remote_ghosts)
{
  return guarded_host([&] {
    if (!h)
      throw ReadError("NULL handle");
    if (remote_ptr) *remote_ptr = h->remote_ptr.data();
    if (remote_ghosts) *remote_ghosts = h->remote_ghosts.data();
  });
}

int spmvx_petsc_read_matrix(const char* path, int32_t nranks, int32_t rank, int32_t dtype,
                            spmvx_matrix* A, spmvx_l2gmap* col_map, spmvx_l2gmap* row_map)
{
  if (A) *A = nullptr;
  if (col_map) *col_map = nullptr;
  if (row_map) *row_map = nullptr;
  spmvx_petsc_csr h = nullptr;
  int st = spmvx_petsc_csr_read(path, nranks, rank, &h);
  if (st != SPMVX_OK)
    return st;
  std::unique_ptr<spmvx_petsc_csr_s> hold(h);
  if (!A || !col_map)
  {
    spmvx::set_last_error("out is NULL");
    return SPMVX_ERR_ARG;
  }
  const int64_t nloc_r = h->row_ranges[rank + 1] - h->row_ranges[rank];
  const int64_t nloc_c = h->col_ranges[rank + 1] - h->col_ranges[rank];
  std::vector<int64_t> csize(nranks), rsize(nranks);
  for (int q = 0; q < nranks; ++q)
  {
    csize[q] = h->col_ranges[q + 1] - h->col_ranges[q];
    rsize[q] = h->row_ranges[q + 1] - h->row_ranges[q];
  }
  st = spmvx_l2gmap_create(nranks, rank, csize.data(), h->ghosts.data(),
                           static_cast<int64_t>(h->ghosts.size()), h->remote_ptr.data(),
                           h->remote_ghosts.data(), col_map);
  if (st == SPMVX_OK && row_map)
    st = spmvx_l2gmap_create(nranks, rank, rsize.data(), nullptr, 0, nullptr, nullptr, row_map);
  if (st == SPMVX_OK)
    st = spmvx_matrix_create(dtype, nloc_r, nloc_c + static_cast<int64_t>(h->ghosts.size()),
                             static_cast<int64_t>(h->colidx.size()), h->rowptr.data(),
                             h->colidx.data(), h->values.data(), A);
  if (st != SPMVX_OK)
  {
    const std::string err = spmvx_last_error();
    if (*col_map) spmvx_l2gmap_destroy(*col_map), *col_map = nullptr;
    if (row_map && *row_map) spmvx_l2gmap_destroy(*row_map), *row_map = nullptr;
    spmvx::set_last_error(err);
  }
  return st;
}

int spmvx_petsc_read_vector(const char* path, int32_t nranks, int32_t rank, double* out,
                            int64_t capacity, int64_t* n_local, int64_t* n_global)
{
  return guarded_host([&] {
    if (nranks < 1 || rank < 0 || rank >= nranks)
      throw ReadError("invalid rank / nranks");
    BigEndianFile f(path);
    if (f.size() < 8)
      throw ReadError("PETSc Vector file is truncated");
    int32_t head[2] = {0, 0};
    f.read_i32(head, 2, "PETSc Vector file");
    if (head[0] != 1211214) // read_petsc.cpp:220-221
      throw ReadError("Bad signature in PETSc Vector file");
    const int64_t n = head[1];
    if (n < 0)
      throw ReadError("PETSc Vector file has a negative size");
    if (f.size() < 8 + 8 * n)
      throw ReadError("PETSc Vector file is truncated");
    const std::vector<int64_t> r = owner_ranges(nranks, n);
    const int64_t nl = r[rank + 1] - r[rank];
    if (n_local) *n_local = nl;
    if (n_global) *n_global = n;
    if (!out)
      return; // size query
    if (capacity < nl)
      throw ReadError("output buffer too small for the local part of the vector");
    f.seek(8 + 8 * r[rank]);
    f.read_f64(out, nl, "PETSc Vector file");
  });
}

} // extern "C"
