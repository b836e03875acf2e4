// Conjugate gradient on the device (spmv/cg.cpp:11-83) and the
// mixed-precision iterative refinement built on it.
//
// One CG iteration = (ghost update) + 5 kernels, all scalars on the device:
//   csr_stream<kDot> / sliced_spmv<kDot>  y = A p, partial p.y   cg.cpp:55-61
//   cg_alpha_finalize  alpha = rnorm_old / p.y                   cg.cpp:62
//   cg_update_r        r -= alpha y, partial ||r||^2             cg.cpp:66,69
//   cg_beta_finalize   ||r||^2, beta, k, converged               cg.cpp:70-79
//   cg_update_xp       x += alpha p, p = beta p + r              cg.cpp:65,76
// Every kernel returns immediately once the device-side `stopped` flag is
// set, so the host replays a captured graph of a fixed number of iterations
// and polls the flag between replays; the iteration count is exact.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "handles.hpp"

namespace spmvx
{
namespace
{

template <typename X, typename Acc>
void ensure_ws(CgWorkspace& w, const Pattern& P, int64_t ncols, int64_t nblocks)
{
  const int64_t n = P.nrows;
  if (w.mem.ptr && w.n == n && w.ncols == ncols && w.nblocks == nblocks && w.elem == sizeof(X))
    return;
  const int vgrid = vec_grid(n > ncols ? n : ncols);
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t vec_n = al(sizeof(X) * static_cast<size_t>(n > 0 ? n : 1));
  const size_t vec_c = al(sizeof(X) * static_cast<size_t>(ncols > 0 ? ncols : 1));
  const int64_t nb = nblocks > 0 ? nblocks : 1;
  const size_t red_s = al(sizeof(Acc) * static_cast<size_t>(nb));
  const size_t red_v = al(sizeof(Acc) * static_cast<size_t>(vgrid));
  const size_t st_s = al(sizeof(CgState<Acc>));
  const size_t total = 2 * vec_n + 3 * vec_c + red_s + 2 * red_v + st_s;
  w.mem.release();
  char* base = static_cast<char*>(w.mem.ensure(total));
  size_t off = 0;
  auto take = [&](size_t b) {
    void* p = base + off;
    off += b;
    return p;
  };
  w.r = take(vec_n);
  w.y = take(vec_n);
  w.p = take(vec_c);
  w.aux = take(vec_c);
  w.aux2 = take(vec_c);
  w.red_spmv = take(red_s);
  w.red_vec0 = take(red_v);
  w.red_vec1 = take(red_v);
  w.state = take(st_s);
  w.n = n;
  w.ncols = ncols;
  w.nblocks = nblocks;
  w.elem = sizeof(X);
  w.vgrid = vgrid;
}

template <typename Acc>
__global__ void cg_state_init(CgState<Acc>* st, int32_t kmax, Acc rtol2)
{
  CgState<Acc> s{};
  s.kmax = kmax;
  s.rtol2 = rtol2;
  s.stopped = 1;
  *st = s;
}

constexpr int kGraphChunk = 16; // iterations per captured graph
// Workgroups of the residual update (and so partial sums for the beta
// finalize): 1024 are enough for bandwidth, and the finalize then reads a
// quarter of the partials of one workgroup per 256 elements (FP64 CG 0.8%
// faster than 4096 on 16.8M rows).
constexpr int kResidGrid = 1024;
constexpr int kTailChunk = 4;   // iterations per graph near convergence

// Whether the iterations still needed, extrapolated from the residual
// decrease between two polls (k1, rho1) and the latest state, end within
// the replay in flight plus one more full chunk: then the next replays use
// the short graph, so that fewer no-op iterations follow convergence.
template <typename Acc>
bool near_convergence(const CgState<Acc>& o, int32_t k1, double rho1, int32_t in_flight)
{
  const double rho = static_cast<double>(o.rnorm_old);
  if (k1 < 0 || o.k <= k1 || !(rho > 0.0) || !(rho < rho1))
    return false;
  const double rate = std::log(rho / rho1) / static_cast<double>(o.k - k1); // < 0
  const double need = std::log(static_cast<double>(o.rtol2) * static_cast<double>(o.rnorm0) / rho);
  return need / rate <= static_cast<double>(in_flight + kGraphChunk);
}

// Host driver of one CG solve.  x has ncols entries; b has nrows.
// Chunks of kGraphChunk and kTailChunk iterations are captured once (per
// matrix, x, history buffer and ghost map) and replayed; the kernels are
// gated by the device state, so replays past convergence are no-ops.  Two
// replays are kept in flight while the host polls the state of the older
// one, switching to the short chunk when the residual decrease so far says
// that convergence is near.
template <typename V, typename X, typename Acc>
int run_cg(spmvx_matrix_s* A, CgWorkspace& w, spmvx_l2gmap_s* col_map, int32_t vec_dtype,
           const X* b, X* x, int32_t kmax, Acc rtol, Acc* hist_dev, hipStream_t s)
{
  Pattern& P = *A->pat;
  const CsrView rb = csr_view(P, A->variant);
  ensure_ws<X, Acc>(w, P, P.ncols, rb.nblocks);
  const int64_t n = P.nrows, ncols = P.ncols;
  auto* st = static_cast<CgState<Acc>*>(w.state);
  X* r = static_cast<X*>(w.r);
  X* y = static_cast<X*>(w.y);
  X* p = static_cast<X*>(w.p);
  Acc* part_s = static_cast<Acc*>(w.red_spmv);
  Acc* part_0 = static_cast<Acc*>(w.red_vec0);
  Acc* part_1 = static_cast<Acc*>(w.red_vec1);
  const int vg = vec_grid(n);
  const int vgr = vg < kResidGrid ? vg : kResidGrid;
  const int g0 = vec_grid(ncols > n ? ncols : n);
  if (!w.stream)
  {
    SPMVX_HIP(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking));
    SPMVX_HIP(hipEventCreateWithFlags(&w.ev_in, hipEventDisableTiming));
    SPMVX_HIP(hipEventCreateWithFlags(&w.ev_poll[0], hipEventDisableTiming));
    SPMVX_HIP(hipEventCreateWithFlags(&w.ev_poll[1], hipEventDisableTiming));
  }
  hipStream_t q = w.stream;
  if (col_map)
  {
    l2gmap_prepare_device(col_map);
    // checked here, not first inside the captured iterations
    SPMVX_REQUIRE(col_map->halo_dtype < 0 || col_map->halo_dtype == vec_dtype,
                  "halo dtype differs from the vector dtype");
  }
  const bool sliced = A->variant == kSlicedVariant;
  const SlicedView sv = sliced ? sliced_view(A) : SlicedView{};
  const void* vals = sliced ? A->svalues.ptr : A->values.ptr;

  // one CG iteration (cg.cpp:52-80); every kernel is gated by the state
  auto iteration = [&](hipStream_t qq) {
    if (col_map)
      l2gmap_update_async(col_map, vec_dtype, p, qq); // cg.cpp:55
    const int64_t nparts =
        sliced ? launch_sliced<V, X, Acc, kDot>(sv, qq, static_cast<const V*>(vals), p, y,
                                                nullptr, part_s, st, 1)
               : launch_csr_stream<V, X, Acc, kDot>(A->variant, rb, qq,
                                                    static_cast<const V*>(vals), p, y, nullptr,
                                                    part_s, st, 1, A->long_rows);
    cg_alpha_finalize<Acc><<<1, kFinBS, 0, qq>>>(part_s, nparts, st, 1);
    cg_update_r<X, Acc><<<vgr, kVecBS, 0, qq>>>(n, r, y, part_1, st, 1);
    cg_beta_finalize<Acc><<<1, kFinBS, 0, qq>>>(part_1, vgr, st, hist_dev, 1);
    cg_update_xp<X, Acc><<<vg, kVecBS, 0, qq>>>(n, x, p, r, st, 1);
  };

  const void* halo = col_map ? col_map->d_halo.ptr : nullptr;
  const int32_t halo_dt = col_map ? col_map->halo_dtype : -2;
  const uint64_t map_id = col_map ? col_map->id : 0;
  if (!w.exec || !w.exec_tail || w.g_x != x || w.g_hist != hist_dev || w.g_map != map_id
      || w.g_halo != halo || w.g_halo_dtype != halo_dt || w.g_variant != A->variant
      || w.g_long_rows != A->long_rows || w.g_mem != w.mem.ptr || w.g_values != vals)
  {
    for (hipGraphExec_t* e : {&w.exec, &w.exec_tail})
    {
      if (*e)
        SPMVX_HIP(hipGraphExecDestroy(*e));
      *e = nullptr;
    }
    for (hipGraph_t* g : {&w.graph, &w.graph_tail})
    {
      if (*g)
        SPMVX_HIP(hipGraphDestroy(*g));
      *g = nullptr;
    }
    w.graph = capture_graph(q, [&] {
      for (int j = 0; j < kGraphChunk; ++j)
        iteration(q);
    });
    check_launch("cg iteration capture");
    SPMVX_HIP(hipGraphInstantiate(&w.exec, w.graph, nullptr, nullptr, 0));
    w.graph_tail = capture_graph(q, [&] {
      for (int j = 0; j < kTailChunk; ++j)
        iteration(q);
    });
    check_launch("cg iteration capture");
    SPMVX_HIP(hipGraphInstantiate(&w.exec_tail, w.graph_tail, nullptr, nullptr, 0));
    w.g_x = x;
    w.g_hist = hist_dev;
    w.g_map = map_id;
    w.g_halo = halo;
    w.g_halo_dtype = halo_dt;
    w.g_variant = A->variant;
    w.g_long_rows = A->long_rows;
    w.g_mem = w.mem.ptr;
    w.g_values = vals;
  }

  // the caller's stream first (b and x may have just been written on it)
  SPMVX_HIP(hipEventRecord(w.ev_in, s));
  SPMVX_HIP(hipStreamWaitEvent(q, w.ev_in, 0));
  cg_state_init<Acc><<<1, 1, 0, q>>>(st, kmax, rtol * rtol);
  cg_init<X, Acc><<<g0, kVecBS, 0, q>>>(n, ncols, b, x, r, p, part_0);
  cg_init_finalize<Acc><<<1, kFinBS, 0, q>>>(part_0, g0, st, hist_dev);
  check_launch("cg_init");
  auto* poll = static_cast<CgState<Acc>*>(w.poll.ensure(2 * sizeof(CgState<double>)));
  int32_t launched = 0;
  int slot = 0, inflight = 0;
  int32_t chunk[2] = {0, 0};
  bool stopped = kmax <= 0, tail = false;
  int32_t k1 = -1; // iteration and residual of the previous poll
  double rho1 = 0.0;
  while (!stopped && launched < kmax)
  {
    chunk[slot] = tail ? kTailChunk : kGraphChunk;
    SPMVX_HIP(hipGraphLaunch(tail ? w.exec_tail : w.exec, q));
    launched += chunk[slot];
    SPMVX_HIP(hipMemcpyAsync(&poll[slot], st, sizeof(CgState<Acc>), hipMemcpyDeviceToHost, q));
    SPMVX_HIP(hipEventRecord(w.ev_poll[slot], q));
    slot ^= 1;
    if (++inflight == 2)
    {
      SPMVX_HIP(hipEventSynchronize(w.ev_poll[slot])); // the older replay
      const CgState<Acc> o = poll[slot];
      stopped = o.stopped != 0;
      inflight = 1;
      tail = near_convergence(o, k1, rho1, chunk[slot ^ 1]);
      k1 = o.k;
      rho1 = static_cast<double>(o.rnorm_old);
    }
  }
  CgState<Acc> fin{};
  SPMVX_HIP(hipMemcpyAsync(A->host_flag.ensure(sizeof(CgState<Acc>)), st, sizeof(CgState<Acc>),
                           hipMemcpyDeviceToHost, q));
  SPMVX_HIP(hipStreamSynchronize(q));
  std::memcpy(&fin, A->host_flag.ptr, sizeof(fin));
  (void)s;
  return fin.k;
}

// ---------------------------------------------------------------- refinement
// x = 0 (FP64)
__global__ void zero_f64(double* x, int64_t n)
{
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    x[i] = 0.0;
}

// r32 = (float)(r64 * scale)
__global__ void scale_to_f32(const double* __restrict__ r, float* __restrict__ r32, int64_t n,
                             const CgState<double>* st)
{
  const double scale = 1.0 / sqrt(st->last);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    r32[i] = static_cast<float>(r[i] * scale);
}

// x64 += norm * (double)d32
__global__ void add_correction(double* __restrict__ x, const float* __restrict__ d, int64_t n,
                               const CgState<double>* st)
{
  const double norm = sqrt(st->last);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n; i += stride)
    x[i] += norm * static_cast<double>(d[i]);
}

// ---------------------------------------------------------------- no-allocation CG
// Scratch of spmvx_cg_noalloc: a file-scope device array, one copy on every
// GPU the library runs on, held by the loaded code object until the process
// ends.  In host code `scratch` names a host-side stand-in, not device
// memory: its address on the current device comes from hipGetSymbolAddress,
// looked up once per device and kept in scratch_addr.
constexpr size_t kNoAllocBytes = SPMVX_CG_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device.
unsigned char* noalloc_scratch()
{
  int dev = 0;
  SPMVX_HIP(hipGetDevice(&dev));
  SPMVX_REQUIRE(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_addr[dev])
  {
    void* p = nullptr;
    SPMVX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SPMVX_REQUIRE(p != nullptr, "no device address for the CG scratch");
    scratch_addr[dev] = p;
  }
  return static_cast<unsigned char*>(scratch_addr[dev]);
}

// Offsets of the CG buffers inside the scratch (256-byte aligned).
struct NoAllocLayout
{
  size_t r = 0, y = 0, p = 0, part_s = 0, part_0 = 0, part_1 = 0, state = 0, hist = 0, total = 0;
};

NoAllocLayout noalloc_layout(int64_t n, int64_t ncols, int64_t nparts, int vgrid,
                             int64_t hist_n)
{
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  NoAllocLayout L;
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t o = off;
    off += al(b);
    return o;
  };
  L.r = take(8 * static_cast<size_t>(n > 0 ? n : 1));
  L.y = take(8 * static_cast<size_t>(n > 0 ? n : 1));
  L.p = take(8 * static_cast<size_t>(ncols > 0 ? ncols : 1));
  L.part_s = take(8 * static_cast<size_t>(nparts > 0 ? nparts : 1));
  L.part_0 = take(8 * static_cast<size_t>(vgrid));
  L.part_1 = take(8 * static_cast<size_t>(vgrid));
  L.state = take(sizeof(CgState<double>));
  L.hist = take(8 * static_cast<size_t>(hist_n > 0 ? hist_n : 1));
  L.total = off;
  return L;
}

// FP64 CG with every buffer in the scratch: the kernels of run_cg, launched
// directly on the caller's stream (no captured graph, no workspace), the stop
// flag read after every chunk of iterations with a copy into host memory of
// the caller's stack.
int run_cg_noalloc(spmvx_matrix_s* A, spmvx_l2gmap_s* col_map, const double* b, double* x,
                   int32_t kmax, double rtol, double* hist_host, hipStream_t s)
{
  Pattern& P = *A->pat;
  const bool sliced = A->variant == kSlicedVariant;
  const CsrView rbv = csr_view_built(P, A->variant);
  const CsrView* rb = rbv.blocks ? &rbv : nullptr;
  SPMVX_REQUIRE(sliced || rb != nullptr,
                "the row blocks of the matrix's kernel variant are not built");
  SPMVX_REQUIRE(!sliced || A->svalues_ready, "the sliced layout of the matrix is not built");
  if (col_map)
  {
    const bool needs_plan = !col_map->indexbuf.empty() || !col_map->ghosts.empty();
    SPMVX_REQUIRE(!needs_plan || col_map->device_ready,
                  "the column map has no device plan yet: call spmvx_l2gmap_update (or "
                  "spmvx_cg) with it once first");
    SPMVX_REQUIRE(col_map->halo_dtype < 0 || col_map->halo_dtype == SPMVX_F64,
                  "halo dtype differs from the vector dtype");
  }
  const int64_t n = P.nrows, ncols = P.ncols;
  const int vg = vec_grid(n);
  const int vgr = vg < kResidGrid ? vg : kResidGrid;
  const int g0 = vec_grid(ncols > n ? ncols : n);
  const int64_t hist_n = hist_host ? static_cast<int64_t>(kmax) + 1 : 0;
  const SlicedView sv = sliced ? sliced_view(A) : SlicedView{}; // built: no allocation
  const int64_t spw = (kSliceBS / kSlice) * (sliced ? sv.rpl : 1);
  const int64_t nparts_max = sliced ? (sv.nslices + spw - 1) / spw : rb->nblocks;
  const NoAllocLayout L = noalloc_layout(n, ncols, nparts_max, g0, hist_n);
  if (L.total > kNoAllocBytes)
    throw Error(SPMVX_ERR_ARG, "spmvx_cg_noalloc: the solve needs " + std::to_string(L.total)
                                   + " bytes of scratch, more than its "
                                   + std::to_string(kNoAllocBytes) + " bytes");
  unsigned char* base = noalloc_scratch();
  double* r = reinterpret_cast<double*>(base + L.r);
  double* y = reinterpret_cast<double*>(base + L.y);
  double* p = reinterpret_cast<double*>(base + L.p);
  double* part_s = reinterpret_cast<double*>(base + L.part_s);
  double* part_0 = reinterpret_cast<double*>(base + L.part_0);
  double* part_1 = reinterpret_cast<double*>(base + L.part_1);
  auto* st = reinterpret_cast<CgState<double>*>(base + L.state);
  double* hist_dev = hist_host ? reinterpret_cast<double*>(base + L.hist) : nullptr;
  const void* vals = sliced ? A->svalues.ptr : A->values.ptr;

  // the iteration of run_cg (cg.cpp:52-80), every kernel gated by the state
  auto iteration = [&]() {
    if (col_map)
      l2gmap_update_async(col_map, SPMVX_F64, p, s); // cg.cpp:55
    const int64_t nparts =
        sliced ? launch_sliced<double, double, double, kDot>(
                     sv, s, static_cast<const double*>(vals), p, y, nullptr, part_s, st, 1)
               : launch_csr_stream<double, double, double, kDot>(
                     A->variant, *rb, s, static_cast<const double*>(vals), p, y, nullptr, part_s,
                     st, 1, A->long_rows);
    cg_alpha_finalize<double><<<1, kFinBS, 0, s>>>(part_s, nparts, st, 1);
    cg_update_r<double, double><<<vgr, kVecBS, 0, s>>>(n, r, y, part_1, st, 1);
    cg_beta_finalize<double><<<1, kFinBS, 0, s>>>(part_1, vgr, st, hist_dev, 1);
    cg_update_xp<double, double><<<vg, kVecBS, 0, s>>>(n, x, p, r, st, 1);
  };

  cg_state_init<double><<<1, 1, 0, s>>>(st, kmax, rtol * rtol);
  cg_init<double, double><<<g0, kVecBS, 0, s>>>(n, ncols, b, x, r, p, part_0);
  cg_init_finalize<double><<<1, kFinBS, 0, s>>>(part_0, g0, st, hist_dev);
  check_launch("cg_init");
  int32_t launched = 0;
  int32_t stopped = kmax <= 0 ? 1 : 0;
  while (!stopped && launched < kmax)
  {
    for (int j = 0; j < kGraphChunk; ++j)
      iteration();
    check_launch("cg iteration");
    launched += kGraphChunk;
    SPMVX_HIP(hipMemcpyAsync(&stopped, &st->stopped, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SPMVX_HIP(hipStreamSynchronize(s));
  }
  CgState<double> fin{};
  SPMVX_HIP(hipMemcpyAsync(&fin, st, sizeof(fin), hipMemcpyDeviceToHost, s));
  SPMVX_HIP(hipStreamSynchronize(s));
  if (hist_host)
  {
    SPMVX_HIP(hipMemcpyAsync(hist_host, hist_dev, sizeof(double) * (static_cast<size_t>(fin.k) + 1),
                             hipMemcpyDeviceToHost, s));
    SPMVX_HIP(hipStreamSynchronize(s));
  }
  return fin.k;
}

// Argument checks shared by spmvx_cg and spmvx_cg_noalloc.
void check_cg_args(spmvx_matrix A, spmvx_l2gmap col_map, const double* b, const double* x,
                   int32_t max_its)
{
  SPMVX_REQUIRE(A != nullptr, "NULL matrix");
  SPMVX_REQUIRE(A->dtype == SPMVX_F64, "spmv::cg needs an FP64 matrix");
  const Pattern& P = *A->pat;
  if (col_map)
  {
    SPMVX_REQUIRE(col_map->local_size == P.nrows, "spmv::cg - Error: b.rows() != A.rows()");
    SPMVX_REQUIRE(col_map->local_size + static_cast<int64_t>(col_map->ghosts.size()) == P.ncols,
                  "column map size differs from the matrix column count");
  }
  else
    SPMVX_REQUIRE(P.ncols == P.nrows, "a matrix with ghost columns needs its column map");
  SPMVX_REQUIRE(max_its >= 0, "max_its must be >= 0");
  SPMVX_REQUIRE((b && x) || P.nrows == 0, "NULL vector");
}

} // namespace
} // namespace spmvx

using namespace spmvx;

extern "C" {

int spmvx_cg(spmvx_matrix A, spmvx_l2gmap col_map, const double* b, double* x, int32_t max_its,
             double rtol, int32_t* its, double* hist_host, spmvx_stream stream)
{
  return guarded([&] {
    check_cg_args(A, col_map, b, x, max_its);
    hipStream_t s = as_stream(stream);
    DevBuf hist;
    double* hd = nullptr;
    if (hist_host)
    {
      hd = static_cast<double*>(hist.ensure(sizeof(double) * (static_cast<size_t>(max_its) + 1)));
    }
    const int k = run_cg<double, double, double>(A, A->ws64, col_map, SPMVX_F64, b, x, max_its,
                                                 rtol, hd, s);
    if (its)
      *its = k;
    if (hist_host)
      SPMVX_HIP(hipMemcpy(hist_host, hd, sizeof(double) * (static_cast<size_t>(k) + 1),
                          hipMemcpyDeviceToHost));
  });
}

int spmvx_cg_noalloc(spmvx_matrix A, spmvx_l2gmap col_map, const double* b, double* x,
                     int32_t max_its, double rtol, int32_t* its, double* hist_host,
                     spmvx_stream stream)
{
  return guarded([&] {
    check_cg_args(A, col_map, b, x, max_its);
    const int k = run_cg_noalloc(A, col_map, b, x, max_its, rtol, hist_host, as_stream(stream));
    if (its)
      *its = k;
  });
}

int spmvx_cg_f32(spmvx_matrix A, const float* b, float* x, int32_t max_its, double rtol,
                 int32_t* its, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(A != nullptr, "NULL matrix");
    SPMVX_REQUIRE(A->dtype == SPMVX_F32 || A->dtype == SPMVX_BF16,
                  "cg_f32 needs an F32 or BF16 matrix");
    const Pattern& P = *A->pat;
    SPMVX_REQUIRE(P.ncols == P.nrows, "cg_f32 needs a square matrix without ghost columns");
    SPMVX_REQUIRE(max_its >= 0, "max_its must be >= 0");
    SPMVX_REQUIRE((b && x) || P.nrows == 0, "NULL vector");
    hipStream_t s = as_stream(stream);
    int k;
    if (A->dtype == SPMVX_F32)
      k = run_cg<float, float, float>(A, A->ws32, nullptr, SPMVX_F32, b, x, max_its,
                                      static_cast<float>(rtol), nullptr, s);
    else
      k = run_cg<bf16_t, float, float>(A, A->ws32, nullptr, SPMVX_F32, b, x, max_its,
                                       static_cast<float>(rtol), nullptr, s);
    if (its)
      *its = k;
  });
}

int spmvx_ir_cg(spmvx_matrix A64, spmvx_matrix Alow, const double* b, double* x,
                int32_t max_outer, double rtol, int32_t inner_max_its, double inner_rtol,
                int32_t flags, spmvx_ir_stats* stats, spmvx_stream stream)
{
  return guarded([&] {
    SPMVX_REQUIRE(A64 && Alow, "NULL matrix");
    SPMVX_REQUIRE(A64->dtype == SPMVX_F64, "outer matrix must be FP64");
    SPMVX_REQUIRE(Alow->dtype == SPMVX_F32 || Alow->dtype == SPMVX_BF16,
                  "inner matrix must be F32 or BF16");
    SPMVX_REQUIRE(Alow->pat == A64->pat, "inner matrix must share the pattern (create_like)");
    Pattern& P = *A64->pat;
    SPMVX_REQUIRE(P.nrows == P.ncols, "refinement needs a square matrix");
    SPMVX_REQUIRE(max_outer >= 0 && max_outer <= 64, "max_outer must be in [0, 64]");
    SPMVX_REQUIRE((b && x) || P.nrows == 0, "NULL vector");
    SPMVX_REQUIRE(inner_max_its >= 0, "inner_max_its must be >= 0");
    SPMVX_REQUIRE((flags & ~SPMVX_IR_ADAPTIVE_INNER) == 0, "unknown refinement flags");
    hipStream_t s = as_stream(stream);
    const int64_t n = P.nrows;
    const CsrView rb64 = csr_view(P, A64->variant);
    ensure_ws<double, double>(A64->ws64, P, P.ncols, rb64.nblocks);
    const CsrView rb32 = csr_view(P, Alow->variant);
    ensure_ws<float, float>(Alow->ws32, P, P.ncols, rb32.nblocks);
    CgWorkspace& w = A64->ws64;
    auto* st = static_cast<CgState<double>*>(w.state);
    double* r = static_cast<double*>(w.r);
    float* r32 = static_cast<float*>(Alow->ws32.aux);
    float* d32 = static_cast<float*>(Alow->ws32.aux2);
    const int vg = vec_grid(n);
    zero_f64<<<vg, kVecBS, 0, s>>>(x, n);
    spmvx_ir_stats local{};
    auto* flag = static_cast<CgState<double>*>(A64->host_flag.ensure(sizeof(CgState<double>)));
    double bnorm = 0.0;
    int32_t outer = 0;
    for (;;)
    {
      // r = b - A x and ||r||^2 in FP64
      const int64_t nparts =
          A64->variant == kSlicedVariant
              ? launch_sliced<double, double, double, kResid>(
                    sliced_view(A64), s, A64->svalues.as<double>(), x, r, b,
                    static_cast<double*>(w.red_spmv), st, 0)
              : launch_csr_stream<double, double, double, kResid>(
                    A64->variant, rb64, s, static_cast<const double*>(A64->values.ptr), x, r, b,
                    static_cast<double*>(w.red_spmv), st, 0, A64->long_rows);
      sum_finalize<double><<<1, kFinBS, 0, s>>>(static_cast<double*>(w.red_spmv), nparts, st);
      check_launch("csr_stream<kResid>");
      SPMVX_HIP(hipMemcpyAsync(flag, st, sizeof(CgState<double>), hipMemcpyDeviceToHost, s));
      SPMVX_HIP(hipStreamSynchronize(s));
      const double rn = std::sqrt(flag->last);
      if (outer == 0)
        bnorm = rn;
      const double rel = bnorm > 0 ? rn / bnorm : 0.0;
      if (outer < 64)
        local.outer_hist[outer] = rel;
      local.rel_residual = rel;
      if (rel <= rtol || outer >= max_outer || rn == 0.0)
        break;
      // inner FP32 solve on the normalised residual; with the adaptive flag
      // an inner solve needs only the reduction that brings the outer
      // residual to half of rtol (never tighter than inner_rtol)
      double irtol = inner_rtol;
      if (flags & SPMVX_IR_ADAPTIVE_INNER)
        irtol = std::max(inner_rtol, 0.5 * rtol / rel);
      scale_to_f32<<<vg, kVecBS, 0, s>>>(r, r32, n, st);
      check_launch("scale_to_f32");
      int k;
      if (Alow->dtype == SPMVX_F32)
        k = run_cg<float, float, float>(Alow, Alow->ws32, nullptr, SPMVX_F32, r32, d32,
                                        inner_max_its, static_cast<float>(irtol), nullptr, s);
      else
        k = run_cg<bf16_t, float, float>(Alow, Alow->ws32, nullptr, SPMVX_F32, r32, d32,
                                         inner_max_its, static_cast<float>(irtol), nullptr, s);
      local.inner_hist[outer] = k;
      local.inner_its += k;
      add_correction<<<vg, kVecBS, 0, s>>>(x, d32, n, st);
      check_launch("add_correction");
      ++outer;
    }
    local.outer_its = outer;
    if (stats)
      *stats = local;
  });
}

} // extern "C"
