// Internal definitions of the opaque handles of the C ABI.
#pragma once

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace spmvx
{

// Device allocation owned by a handle.
struct DevBuf
{
  void* ptr = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release()
  {
    if (ptr)
      (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  // Grows (never shrinks) to n bytes; new memory is zeroed.
  void* ensure(size_t n)
  {
    if (n > bytes)
    {
      release();
      SPMVX_HIP(hipMalloc(&ptr, n));
      SPMVX_HIP(hipMemset(ptr, 0, n));
      bytes = n;
    }
    return ptr;
  }
  template <typename T>
  T* as() const
  {
    return static_cast<T*>(ptr);
  }
};

// Pinned host allocation (hipHostMalloc).
struct PinnedBuf
{
  void* ptr = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf()
  {
    if (ptr)
      (void)hipHostFree(ptr);
  }
  void* ensure(size_t n)
  {
    if (n > bytes)
    {
      if (ptr)
        (void)hipHostFree(ptr);
      ptr = nullptr;
      SPMVX_HIP(hipHostMalloc(&ptr, n, hipHostMallocDefault));
      bytes = n;
    }
    return ptr;
  }
};

// Row-binning of a pattern for one staging chunk size: block descriptors
// (first row, first entry), nblocks + 1 of them; and, built on request for
// the short-index kernels, every entry's 16-bit column index and every
// block's four window bases (see csr_stream).
struct RowBlocks
{
  DevBuf desc;
  int64_t nblocks = 0;
  bool short_built = false;
  int short_kind = 0; // 0: 32-bit only, 1: one base per block, 2: four windows
  DevBuf col16, wbase;
};

// CSR pattern + row-binning analysis on the device; shared between a matrix
// and its reduced-precision copies.  The host row pointer is kept so that
// blocks for another chunk size can be built on demand.
// Row patterns of a CSR pattern (sliced layout): the 16-bit pattern id of
// every row and the pattern table (length, offsets col - row) in rows of
// wmax + 1 entries.
struct RowPatterns
{
  int32_t npat = 0, wmax = 0, dict_n = 0;
  DevBuf pid, dict;
  std::vector<uint16_t> h_pid;
  std::vector<int32_t> h_len; // length of each pattern
};

// Slice descriptors (first value slot / 64, padded width) for one number
// of entries per lane and load (VW).
struct SliceLayout
{
  int64_t nslices = 0, slots = 0;
  DevBuf sdesc;
};

// Limits of the sliced layout: distinct patterns (16-bit ids), the longest
// row, and the pattern table held in LDS (32 KiB).
constexpr int32_t kMaxPatterns = 65535;
constexpr int32_t kMaxSliceWidth = 256;
constexpr int32_t kMaxDictEntries = 8192;

struct Pattern
{
  int64_t nrows = 0, ncols = 0, nnz = 0;
  DevBuf rowptr, colidx;
  std::vector<int32_t> h_rowptr;
  std::map<std::pair<int, int>, std::unique_ptr<RowBlocks>> blocks; // by (chunk, rows)
  int device = 0;
  const RowBlocks& blocks_for(int chunk, int rows = kBS);
  // the blocks for (chunk, rows) if they have been built, else null (never
  // allocates)
  const RowBlocks* find_blocks(int chunk, int rows);
  // sliced layout, analysed on first use; null with a reason if not eligible
  std::unique_ptr<RowPatterns> rpat;
  std::string rpat_error;
  bool rpat_tried = false;
  std::map<int, std::unique_ptr<SliceLayout>> slices; // by VW
  const RowPatterns* row_patterns();
  const SliceLayout& slice_layout(int vw);
  // the lazily built parts above are shared by every matrix made with
  // spmvx_matrix_create_like from one base, which callers may use from
  // different threads
  std::recursive_mutex lazy_mu;
};

// Row blocks for the ordered stream kernel: consecutive rows, at most
// max_rows rows and max_nnz entries per block; a row longer than max_nnz
// is a block of its own.  Returns (first row, first entry) pairs.
std::vector<int2> row_blocks(const int32_t* rowptr, int64_t nrows, int64_t max_nnz,
                             int64_t max_rows);

// Default kernel variant of a value type (overridable with SPMVX_SPMV_VARIANT).
int default_variant(int32_t dtype);

// Row blocks used by the CSR kernel of variant `var`.
inline const RowBlocks& variant_blocks(Pattern& P, int var)
{
  return P.blocks_for(variant_block_nnz(var), variant_block_rows(var));
}
inline const RowBlocks* built_variant_blocks(Pattern& P, int var)
{
  return P.find_blocks(variant_block_nnz(var), variant_block_rows(var));
}

// Builds the short column indices of a row blocking if missing (device
// analysis, synchronous).
void build_short_index(Pattern& P, const RowBlocks& rb, int chunk);

// The arrays a CSR kernel of variant `var` reads (blocks and, for the
// short-index variants, the short indices are built if missing).
CsrView csr_view(Pattern& P, int var);
// The same without building anything: null blocks if they do not exist yet.
CsrView csr_view_built(Pattern& P, int var);

// Builds the sliced layout of a matrix (row patterns, slices, values) if it
// is missing; throws with the reason when the pattern is not eligible.
SlicedView sliced_view(spmvx_matrix_s* A);

inline bool is_complex(int32_t dtype) { return dtype == SPMVX_C64 || dtype == SPMVX_C32; }

// Default kernel variant of a matrix: the short-index counterpart of the
// dtype's default when the matrix fits 16-bit column indices (builds that
// analysis), else the dtype's default.
int matrix_default_variant(spmvx_matrix_s* A);

// Bytes of one stored value of a matrix dtype.
inline size_t dtype_size(int32_t dtype)
{
  switch (dtype)
  {
  case SPMVX_F64: return 8;
  case SPMVX_F32: return 4;
  case SPMVX_BF16: return 2;
  case SPMVX_C64: return 16;
  case SPMVX_C32: return 8;
  default: throw Error(SPMVX_ERR_ARG, "unknown dtype");
  }
}

// Transposed copy of A (transpose.hip), built on stream s, which the call
// synchronises; and the re-gather of its values after a value update
// (enqueued on s).
void build_transpose(spmvx_matrix_s* A, hipStream_t s);
void gather_transposed_values(spmvx_matrix_s* A, hipStream_t s);

// Waits for the last value update of A (spmvx_matrix_set_values runs on the
// caller's stream) before work that reads the values outside that stream.
void wait_values(spmvx_matrix_s* A);

// Records the launches of f on stream q (capture mode thread-local) as a
// graph.  If f throws, the capture is ended and its partial graph dropped
// before the exception propagates, so the stream stays usable.
template <typename F>
hipGraph_t capture_graph(hipStream_t q, F&& f)
{
  SPMVX_HIP(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
  try
  {
    f();
  }
  catch (...)
  {
    hipGraph_t partial = nullptr;
    if (hipStreamEndCapture(q, &partial) == hipSuccess && partial)
      (void)hipGraphDestroy(partial);
    throw;
  }
  hipGraph_t g = nullptr;
  SPMVX_HIP(hipStreamEndCapture(q, &g));
  return g;
}

// Grid of the vector kernels for n elements.
inline int vec_grid(int64_t n)
{
  int64_t g = (n + kVecBS - 1) / kVecBS;
  if (g > 4096)
    g = 4096;
  if (g < 1)
    g = 1;
  return static_cast<int>(g);
}

// Per-precision scratch of the CG solvers, owned by a matrix handle,
// including the captured graph of a chunk of iterations and the stream it
// runs on (graphs cannot be captured on the null stream).
struct CgWorkspace
{
  CgWorkspace() = default;
  CgWorkspace(const CgWorkspace&) = delete;
  CgWorkspace& operator=(const CgWorkspace&) = delete;
  ~CgWorkspace()
  {
    for (hipGraphExec_t e : {exec, exec_tail})
      if (e)
        (void)hipGraphExecDestroy(e);
    for (hipGraph_t g : {graph, graph_tail})
      if (g)
        (void)hipGraphDestroy(g);
    for (hipEvent_t e : {ev_in, ev_poll[0], ev_poll[1]})
      if (e)
        (void)hipEventDestroy(e);
    if (stream)
      (void)hipStreamDestroy(stream);
  }
  hipStream_t stream = nullptr;
  hipEvent_t ev_in = nullptr;
  hipEvent_t ev_poll[2] = {nullptr, nullptr};
  PinnedBuf poll;
  hipGraph_t graph = nullptr; // kGraphChunk iterations
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph_tail = nullptr; // kTailChunk iterations (near convergence)
  hipGraphExec_t exec_tail = nullptr;
  const void* g_x = nullptr;   // arguments the graphs were captured with
  const void* g_hist = nullptr;
  uint64_t g_map = 0; // id of the column map (0: none)
  const void* g_halo = nullptr;
  int32_t g_halo_dtype = -2;
  int32_t g_variant = -1;
  int32_t g_long_rows = -1;
  const void* g_mem = nullptr;
  const void* g_values = nullptr;
  DevBuf mem;
  int64_t n = 0, ncols = 0, nblocks = 0;
  size_t elem = 0;
  // carved pointers
  void* r = nullptr;
  void* y = nullptr;
  void* p = nullptr;
  void* aux = nullptr; // extra vector (refinement: residual / correction)
  void* aux2 = nullptr;
  void* state = nullptr;
  void* red_spmv = nullptr;
  void* red_vec0 = nullptr;
  void* red_vec1 = nullptr;
  int vgrid = 0;
};

} // namespace spmvx

struct spmvx_matrix_s
{
  int32_t dtype = SPMVX_F64;
  int32_t variant = 0;   // SpMV kernel variant (kVariants)
  int32_t long_rows = 0; // 1: rows longer than a block are summed in parallel
  std::shared_ptr<spmvx::Pattern> pat;
  spmvx::DevBuf values;
  // transposed copy for spmv_t (built on first use): tvalues[d] = values[tperm[d]]
  std::shared_ptr<spmvx::Pattern> tpat;
  spmvx::DevBuf tvalues, tperm;
  // recorded by spmvx_matrix_set_values on the caller's stream
  hipEvent_t values_event = nullptr;
  bool values_pending = false;
  // values in the sliced row-pattern layout (built when that kernel is chosen)
  spmvx::DevBuf svalues;
  bool svalues_ready = false;
  // solver scratch (FP64 / FP32)
  spmvx::CgWorkspace ws64, ws32;
  spmvx::PinnedBuf host_flag;
  spmvx_matrix_s() = default;
  spmvx_matrix_s(const spmvx_matrix_s&) = delete;
  spmvx_matrix_s& operator=(const spmvx_matrix_s&) = delete;
  ~spmvx_matrix_s()
  {
    if (values_event)
      (void)hipEventDestroy(values_event);
  }
};

struct spmvx_l2gmap_s
{
  // unique for the process (never reused, unlike the handle's address): keys
  // what a captured CG graph recorded from this map
  uint64_t id = 0;
  int32_t nranks = 1, rank = 0;
  std::vector<int64_t> ranges;
  std::vector<int64_t> ghosts;
  std::vector<int32_t> neighbours, send_count, recv_count, send_offset, recv_offset;
  std::vector<int32_t> indexbuf;
  // deterministic scatter-add plan of reverse_update: unique targets and the
  // positions in indexbuf (ascending) that add into each
  std::vector<int32_t> rev_target, rev_ptr, rev_pos;
  int64_t local_size = 0;
  // device side (allocated on first use)
  spmvx::DevBuf d_indexbuf, d_send, d_rev_target, d_rev_ptr, d_rev_pos;
  spmvx::PinnedBuf h_send, h_in;
  spmvx::DevBuf d_halo;    // device copy of the halo (what the neighbours send)
  int32_t halo_dtype = -1; // dtype of the halo in d_halo, -1 = zeros
  bool device_ready = false;
};

namespace spmvx
{
// Device plan upload (l2gmap.hip) and the update used by the solvers.
void l2gmap_prepare_device(spmvx_l2gmap_s* m);
void l2gmap_update_async(spmvx_l2gmap_s* m, int32_t dtype, void* vec_dev, hipStream_t s);
} // namespace spmvx
