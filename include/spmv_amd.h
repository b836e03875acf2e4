/*
 * spmv_amd.h -- C ABI of the MI355X (gfx950) implementation of libspmv's
 * hot path: CSR SpMV (Matrix<T>::operator*), the ghost update of the
 * local-to-global map (L2GMap::update / reverse_update) and the conjugate
 * gradient solver (spmv::cg), plus the three extensions of this library
 * (mixed-precision refinement, batched persistent CG, adjoint / SDDMM).
 *
 * Conventions
 *   - Every function returns 0 (SPMVX_OK) on success or a positive
 *     SPMVX_ERR_* code; spmvx_last_error() then returns the message (thread
 *     local).  Messages raised where the reference throws use the
 *     reference's text (e.g. "Ghosts must be sorted").
 *   - Pointers named *_dev are device pointers (hipMalloc / torch CUDA
 *     tensors), *_host are host pointers.  Sizes are element counts.
 *   - `stream` is a hipStream_t passed as void* (NULL = null stream).  All
 *     work is enqueued on it; functions that return host results
 *     synchronise it.
 *   - Vector precision follows the matrix value type: SPMVX_F64 matrices
 *     take double vectors; SPMVX_F32 and SPMVX_BF16 matrices take float
 *     vectors (BF16 values are widened to FP32, FP32 accumulation).
 *   - Complex matrices (SPMVX_C64 / SPMVX_C32, the reference's
 *     Matrix<std::complex<T>>, Matrix.cpp:441-444) take complex vectors of
 *     the same precision, (re, im) interleaved; host value arrays are then
 *     2*nnz doubles.  They support the products, transposed products, value
 *     updates and the L2GMap operations (kernel variants 0-4); CG, refinement,
 *     SDDMM, the sliced layout and create_like are for real values.
 *   - Handles are not thread safe; calls on one handle must be serialised,
 *     and so must calls that use one L2GMap (a matrix's column map, a CG's
 *     map) -- as in the reference, where update is collective.  Matrices
 *     made by spmvx_matrix_create_like from one base share its pattern; the
 *     library locks what it builds there, so they may be used from
 *     different threads.
 */
#ifndef SPMV_AMD_H
#define SPMV_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPMVX_ABI_VERSION 2

enum spmvx_status
{
  SPMVX_OK = 0,
  SPMVX_ERR_ARG = 1,         /* invalid argument (reference: runtime_error) */
  SPMVX_ERR_HIP = 2,         /* HIP runtime failure                         */
  SPMVX_ERR_STATE = 3,       /* handle in the wrong state                   */
  SPMVX_ERR_UNSUPPORTED = 4  /* not available in this build                 */
};

/* Scalar types; replaces spmv::mpi_type<T>() (spmv/mpi_type.h:11-33). */
enum spmvx_dtype
{
  SPMVX_F64 = 0,
  SPMVX_F32 = 1,
  SPMVX_BF16 = 2, /* matrix values only */
  SPMVX_C64 = 3,  /* complex double: (re, im) pairs of doubles */
  SPMVX_C32 = 4   /* complex float: (re, im) pairs of floats */
};

typedef struct spmvx_matrix_s* spmvx_matrix;
typedef struct spmvx_l2gmap_s* spmvx_l2gmap;
typedef struct spmvx_batch_s* spmvx_batch;
typedef struct spmvx_petsc_csr_s* spmvx_petsc_csr;
typedef void* spmvx_stream;

/* Last error message of the calling thread ("" if none). */
const char* spmvx_last_error(void);
/* SPMVX_ABI_VERSION of the loaded library. */
int spmvx_abi_version(void);

/* ===================================================================== */
/* L2GMap (spmv/L2GMap.h:29-30, spmv/L2GMap.cpp:15-122)                   */
/* ===================================================================== */
/*
 * Builds the exchange plan of rank `rank` in a decomposition of `nranks`
 * contiguous ownership ranges (local_sizes[q] indices owned by rank q, in
 * rank order: the MPI_Allgather of L2GMap.cpp:25).  One process drives one
 * rank: the other ranks are virtual, and their side of the symmetric
 * neighbour discovery (L2GMap.cpp:57-115) is given explicitly:
 *   ghosts[0..nghosts)           this rank's ghost global indices, sorted
 *                                ascending (L2GMap.h:27-30)
 *   remote_ptr[0..nranks], remote_ghosts[...]
 *                                for every rank q, the sorted global indices
 *                                of q's ghosts that this rank owns
 *                                (remote_ptr may be NULL when there are none)
 * The plan equals the one the reference computes with MPI: neighbours,
 * send/recv counts and offsets (send offsets shifted by local_size,
 * L2GMap.cpp:119-121) and indexbuf (owner-local indices to pack).
 * Errors: "Ghosts must be sorted", "Ghost index in local range",
 * "Ghost index outside the global range".
 */
int spmvx_l2gmap_create(int32_t nranks, int32_t rank, const int64_t* local_sizes,
                        const int64_t* ghosts, int64_t nghosts,
                        const int64_t* remote_ptr, const int64_t* remote_ghosts,
                        spmvx_l2gmap* out);
int spmvx_l2gmap_destroy(spmvx_l2gmap map);

/* local_size / num_ghosts / global_size / global_offset (L2GMap.cpp:187-199),
 * |indexbuf| and the neighbour count.  Any output may be NULL. */
int spmvx_l2gmap_info(spmvx_l2gmap map, int64_t* local_size, int64_t* num_ghosts,
                      int64_t* global_size, int64_t* global_offset,
                      int64_t* num_indices, int32_t* num_neighbours);
/* Copies the plan arrays out (host); sizes from spmvx_l2gmap_info: the count
 * arrays have max(num_neighbours,1) entries, offsets one more
 * (L2GMap.cpp:86-107).  Any output may be NULL. */
int spmvx_l2gmap_plan(spmvx_l2gmap map, int32_t* neighbours, int32_t* send_count,
                      int32_t* recv_count, int32_t* send_offset, int32_t* recv_offset,
                      int32_t* indexbuf);
/* Ghost global indices (L2GMap.h:75). */
int spmvx_l2gmap_ghosts(spmvx_l2gmap map, int64_t* ghosts_host);
/* global_to_local (L2GMap.cpp:172-185); an index neither owned nor ghosted
 * is an error (the reference asserts). */
int spmvx_l2gmap_global_to_local(spmvx_l2gmap map, int64_t global_index, int32_t* local);

/*
 * Single-process transport.  What the neighbours would send into this rank's
 * ghost region comes from a host "halo" buffer of num_ghosts values of
 * `dtype` (F64, F32, C64 or C32), copied into library-owned device memory
 * (one synchronous host-to-device copy per call); NULL means zeros (the
 * default).  Every later update unpacks that device copy, so call this only
 * when no update is in flight (synchronise the stream first), and again
 * whenever the neighbours' values change.
 */
int spmvx_l2gmap_set_halo(spmvx_l2gmap map, int32_t dtype, const void* halo_host);

/*
 * L2GMap::update<T> (L2GMap.cpp:126-150) on a device vector of
 * local_size + num_ghosts values of `dtype`: packs vec[indexbuf[i]] into the
 * device send buffer (what the neighbours would receive), and unpacks the
 * halo into vec[local_size ..).  If send_host is not NULL the packed buffer
 * is also copied there (|indexbuf| values) and the stream is synchronised.
 */
int spmvx_l2gmap_update(spmvx_l2gmap map, int32_t dtype, void* vec_dev, void* send_host,
                        spmvx_stream stream);

/*
 * L2GMap::reverse_update<T> (L2GMap.cpp:152-170): the ghost region
 * vec[local_size ..) is what is sent to the owners (copied to ghost_out_host
 * when not NULL), and incoming_host (|indexbuf| values, what the neighbours
 * send back; NULL = zeros) is scatter-added: vec[indexbuf[i]] += incoming[i],
 * repeated indices accumulating in index order.  Synchronises the stream.
 */
int spmvx_l2gmap_reverse_update(spmvx_l2gmap map, int32_t dtype, void* vec_dev,
                                const void* incoming_host, void* ghost_out_host,
                                spmvx_stream stream);

/* ===================================================================== */
/* Matrix (spmv/Matrix.h:30-53, spmv/Matrix.cpp:13-211)                   */
/* ===================================================================== */
/*
 * Copies a compressed row-major CSR (rowptr[nrows+1], colidx[nnz] with
 * 0 <= colidx < ncols, sorted within each row, values[nnz] as double) from
 * the host to the device, stores the values as `dtype` (F64, F32 or BF16,
 * round to nearest even) and runs the row-binning analysis (the analogue of
 * the inspector call mkl_sparse_optimize, Matrix.cpp:43).  Columns
 * [0, local_size) are owned, [local_size, ncols) ghosts.
 * Replaces Matrix<T>::Matrix (Matrix.cpp:13-22).
 */
int spmvx_matrix_create(int32_t dtype, int64_t nrows, int64_t ncols, int64_t nnz,
                        const int32_t* rowptr_host, const int32_t* colidx_host,
                        const double* values_host, spmvx_matrix* out);
/* A matrix with the same (shared, not copied) pattern as `base` and its
 * values converted to `dtype` on the device: the low-precision copy used by
 * mixed-precision refinement. */
int spmvx_matrix_create_like(spmvx_matrix base, int32_t dtype, spmvx_matrix* out);
int spmvx_matrix_destroy(spmvx_matrix A);
int spmvx_matrix_info(spmvx_matrix A, int32_t* dtype, int64_t* nrows, int64_t* ncols,
                      int64_t* nnz, int64_t* nblocks);
/* Device pointers of the stored arrays (values in the stored dtype). */
int spmvx_matrix_arrays(spmvx_matrix A, const int32_t** rowptr_dev,
                        const int32_t** colidx_dev, const void** values_dev);
/* Selects the SpMV kernel variant of A (staging chunk, loads in flight,
 * streaming loads, index width; 0 <= variant < spmvx_num_kernels(), -1 =
 * the matrix's default).  All variants give identical results; used by
 * tuning.  No reference counterpart (the inspector of Matrix.cpp:43 picks its
 * own kernel).  Variants 8-11 read 16-bit column indices: for every row block
 * whose columns fall into at most four windows of 16384 columns (stencils,
 * banded and locally numbered matrices) the analysis stores each column as a
 * window number and an offset (2 instead of 4 bytes per entry); other blocks
 * keep the 32-bit indices.  A matrix's default is the 16-bit counterpart of
 * its dtype's default kernel when that covers at least half of the entries.
 * The last variant, SPMVX_SLICED_VARIANT, is the sliced row-pattern layout:
 * for matrices whose rows repeat a few column-offset patterns (stencils), a
 * copy of the values is kept in 64-row slices with a 16-bit pattern id per
 * row instead of column indices (the CSR arrays stay as they are).  Selecting
 * it builds that copy; a matrix that does not fit the layout (more than
 * 65535 patterns, rows over 256 entries, a pattern table over 8192 entries)
 * returns SPMVX_ERR_ARG with the reason and keeps its variant.  Transposed
 * products use the default CSR variant when the sliced layout is selected. */
#define SPMVX_SLICED_VARIANT 12
int spmvx_matrix_set_kernel(spmvx_matrix A, int32_t variant);
int spmvx_num_kernels(void);
/* How rows longer than a kernel's staging block (2048-8192 entries, per
 * variant) are summed.  0 (default): in stored order by one lane -- the
 * sequential sum of Matrix.cpp:203, bit-identical, one dependent add per
 * entry.  1: by the whole workgroup, each thread adding a strided subset in
 * order, then a fixed-order tree -- deterministic, rounded differently from
 * the sequential sum (relative to sum |a_ij x_j| the difference is a few
 * ulps), and much faster for rows of 10^4 entries and more.  Real values only.
 * Applies to products, transposed products, CG and refinement with A; copies
 * made by spmvx_matrix_create_like inherit it.  No reference counterpart. */
int spmvx_matrix_set_long_rows(spmvx_matrix A, int32_t parallel);
/* The variant A uses and a short description of its kernel (written into
 * name[0:name_len), NUL-terminated; name may be NULL). */
int spmvx_matrix_get_kernel(spmvx_matrix A, int32_t* variant, char* name, int32_t name_len);
/* Row-pattern analysis of A for the sliced layout (built if needed; no
 * values are copied): number of distinct patterns, longest row and value
 * slots of the layout for A's dtype (nnz plus padding).  SPMVX_ERR_ARG with
 * the reason when A does not fit the layout.  No reference counterpart. */
int spmvx_matrix_sliced_info(spmvx_matrix A, int32_t* npatterns, int32_t* wmax, int64_t* slots);
/* Overwrite the values from nnz doubles on the device (pattern unchanged). */
int spmvx_matrix_set_values(spmvx_matrix A, const double* values_dev, spmvx_stream stream);

/*
 * y[nrows] = A x[ncols] -- Matrix<T>::operator* (Matrix.h:37-39,
 * Matrix.cpp:158-166,199-204).  For SPMVX_F64 every row is summed in stored
 * column order with separately rounded products and sums, starting from 0:
 * the result is bit-identical to the reference's Eigen product.  x and y
 * must not overlap (SPMVX_ERR_ARG).
 */
int spmvx_spmv(spmvx_matrix A, const void* x_dev, void* y_dev, spmvx_stream stream);

/* spmvx_spmv, with two hipEvent_t that record the product kernel's own start
 * and end (hipExtLaunchKernel), so that hipEventElapsedTime between them is
 * the kernel's execution time without the dispatch before it -- the figure a
 * kernel trace reports.  A measurement aid; no reference counterpart. */
int spmvx_spmv_timed(spmvx_matrix A, const void* x_dev, void* y_dev, spmvx_stream stream,
                     void* start_event, void* stop_event);

/* y[ncols] = A^T x[nrows] -- Matrix<T>::transpmult (Matrix.h:41-43,
 * Matrix.cpp:206-211).  Uses an analysis-time transposed copy; every output
 * entry is summed in ascending row order (bit-identical for SPMVX_F64).
 * x and y must not overlap. */
int spmvx_spmv_t(spmvx_matrix A, const void* x_dev, void* y_dev, spmvx_stream stream);

/* ===================================================================== */
/* CG (spmv/cg.h:27-29, spmv/cg.cpp:11-83)                                */
/* ===================================================================== */
/*
 * Unpreconditioned CG with x0 = 0, the reference's loop and stopping rule
 * (||r_k||^2 / ||b||^2 < rtol^2 checked after the direction update, or
 * k == max_its).  A must be SPMVX_F64 with nrows == col_map local_size.
 * col_map (may be NULL when A has no ghost columns) is updated on the
 * direction vector every iteration (cg.cpp:55) through the transport above.
 *   b_dev[nrows], x_dev[ncols] (ghost entries returned as 0), *its = k.
 *   hist_host (may be NULL): max_its+1 doubles, hist[j] = ||r_j||^2.
 * Errors: "spmv::cg - Error: b.rows() != A.rows()" style argument checks.
 */
int spmvx_cg(spmvx_matrix A, spmvx_l2gmap col_map, const double* b_dev, double* x_dev,
             int32_t max_its, double rtol, int32_t* its, double* hist_host,
             spmvx_stream stream);

/*
 * spmvx_cg without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * checks and results as spmvx_cg (the same kernels in the same order: x, *its
 * and the history are bit-identical), but every buffer of the solve (r, A p,
 * the direction vector, partial sums, the device-side scalars and the
 * history) lives in a fixed device array of SPMVX_CG_NOALLOC_SCRATCH_BYTES
 * that the library holds on every GPU it runs on until the process ends, and
 * the kernels are launched directly on `stream` (no captured graph, no
 * internal stream or event).  The call needs about 8 * (2 * nrows + ncols)
 * bytes plus 8 * (max_its + 1) with a history; a solve that needs more is
 * rejected with SPMVX_ERR_ARG before any kernel runs.  Preconditions (else
 * SPMVX_ERR_ARG, nothing launched): the row blocks of A's kernel variant and,
 * for the sliced variant, its layout exist (spmvx_matrix_create and
 * spmvx_matrix_set_kernel build them), and a col_map with ghosts or pack
 * indices has been used once (spmvx_l2gmap_update or spmvx_cg), which builds
 * its device plan.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two host threads, or two streams in flight) race and
 * must be serialised by the caller.  Synchronises `stream`.
 */
#define SPMVX_CG_NOALLOC_SCRATCH_BYTES (32u << 20)
int spmvx_cg_noalloc(spmvx_matrix A, spmvx_l2gmap col_map, const double* b_dev, double* x_dev,
                     int32_t max_its, double rtol, int32_t* its, double* hist_host,
                     spmvx_stream stream);

/* ===================================================================== */
/* PETSc binary input (spmv/read_petsc.h:10-31, spmv/read_petsc.cpp)      */
/* ===================================================================== */
/*
 * Files as PETSc writes them with -ksp_view_mat / -ksp_view_rhs binary:
 * big-endian, 32-bit indices.  Rows (and columns) are split into `nranks`
 * near-equal contiguous ranges, the first (N mod nranks) one longer
 * (owner_ranges, read_petsc.cpp:13-30); rank `rank` gets its rows, its
 * owned columns numbered first and the other referenced columns as ghosts
 * in ascending global order (read_petsc.cpp:115-133), every row sorted by
 * local column as Eigen stores it (read_petsc.cpp:152-160).
 * Errors carry the reference's texts ("Could not open file", "Bad signature
 * in PETSc Matrix file", "Bad signature in PETSc Vector file"), plus
 * truncation / range / duplicate-column checks the reference leaves to
 * assert.
 */
/* Host-side parse only (no GPU needed): local CSR, ghosts and, for the
 * exchange plan, which of this rank's columns every other rank ghosts. */
int spmvx_petsc_csr_read(const char* path, int32_t nranks, int32_t rank, spmvx_petsc_csr* out);
int spmvx_petsc_csr_destroy(spmvx_petsc_csr h);
int spmvx_petsc_csr_info(spmvx_petsc_csr h, int64_t* nrows_global, int64_t* ncols_global,
                         int64_t* nrows_local, int64_t* ncols_local, int64_t* nghosts,
                         int64_t* nnz_local);
/* Host pointers owned by h (valid until spmvx_petsc_csr_destroy). */
int spmvx_petsc_csr_arrays(spmvx_petsc_csr h, const int32_t** rowptr, const int32_t** colidx,
                           const double** values, const int64_t** ghosts);
/* For the exchange plan: remote_ghosts[remote_ptr[q]:remote_ptr[q+1]) are the
 * columns of this rank that rank q ghosts (ascending; empty for q = rank),
 * the remote_* arguments of spmvx_l2gmap_create. */
int spmvx_petsc_csr_remote(spmvx_petsc_csr h, const int64_t** remote_ptr,
                           const int64_t** remote_ghosts);
/* Replaces spmv::read_petsc_binary (read_petsc.cpp:32-191): the matrix (values
 * stored as dtype), its column map (the exchange plan of this rank) and its
 * row map (no ghosts; may be NULL). */
int spmvx_petsc_read_matrix(const char* path, int32_t nranks, int32_t rank, int32_t dtype,
                            spmvx_matrix* A, spmvx_l2gmap* col_map, spmvx_l2gmap* row_map);
/* Replaces spmv::read_petsc_binary_vector (read_petsc.cpp:193-260): this
 * rank's part into out_host[0:n_local) (host memory, capacity entries).
 * out_host = NULL only reports n_local and n_global. */
int spmvx_petsc_read_vector(const char* path, int32_t nranks, int32_t rank, double* out_host,
                            int64_t capacity, int64_t* n_local, int64_t* n_global);

/* ===================================================================== */
/* Extension 1: mixed-precision iterative refinement (no reference        */
/* counterpart; wraps spmvx_cg's loop as the inner solver)                */
/* ===================================================================== */
typedef struct
{
  int32_t outer_its;      /* refinement steps taken                          */
  int32_t inner_its;      /* total inner CG iterations                        */
  double rel_residual;    /* final ||b - A x|| / ||b|| in FP64                */
  double outer_hist[64];  /* ||b - A x_j|| / ||b|| before step j (j<=outer)   */
  int32_t inner_hist[64]; /* inner iterations of step j                       */
} spmvx_ir_stats;

/*
 * Solves A64 x = b to ||b - A64 x|| / ||b|| <= rtol.  Each outer step forms
 * r = b - A64 x (FP64 SpMV) and ||r|| in FP64 on the device, runs CG in FP32
 * on Alow d = r / ||r|| (Alow: SPMVX_F32 or SPMVX_BF16 copy with the same
 * pattern; FP32 vectors and accumulation; stop at relative residual
 * inner_rtol or inner_max_its) and updates x += ||r|| d in FP64.
 * Every inner solve starts from d = 0, p = r (classic refinement).  flags:
 * 0, or SPMVX_IR_ADAPTIVE_INNER: step j's inner solve stops at relative
 * residual max(inner_rtol, rtol / (2 ||r_j|| / ||b||)) -- only the reduction
 * that brings the outer residual to rtol / 2 (the last step usually needs far
 * fewer inner iterations than inner_rtol asks).
 * x_dev[ncols of A64]; ghost columns are treated as zero.
 */
#define SPMVX_IR_ADAPTIVE_INNER 1
int spmvx_ir_cg(spmvx_matrix A64, spmvx_matrix Alow, const double* b_dev, double* x_dev,
                int32_t max_outer, double rtol, int32_t inner_max_its, double inner_rtol,
                int32_t flags, spmvx_ir_stats* stats, spmvx_stream stream);

/* FP32 CG on an SPMVX_F32 / SPMVX_BF16 matrix (the inner solver alone). */
int spmvx_cg_f32(spmvx_matrix A, const float* b_dev, float* x_dev, int32_t max_its,
                 double rtol, int32_t* its, spmvx_stream stream);

/* ===================================================================== */
/* Extension 2: batched persistent CG (no reference counterpart; each     */
/* system runs spmv::cg's loop, cg.cpp:50-80, in FP32)                     */
/* ===================================================================== */
/*
 * nsys independent SPD systems given as the diagonal blocks of one
 * block-diagonal CSR matrix: system s owns rows and columns
 * [sys_row[s], sys_row[s+1]); rowptr[sys_row[nsys]+1] and colidx (global
 * column indices, each inside its own system) and FP32 values.  nrhs (1, 2
 * or 4) right-hand sides per system.  Vectors are [row][nrhs] row-major.
 * With the environment variable SPMVX_BATCH_REORDER=1 at creation, every
 * system is stored in reverse Cuthill-McKee order (b and x stay in the
 * caller's order).
 */
int spmvx_batch_create(int32_t nsys, const int32_t* sys_row_host, const int32_t* rowptr_host,
                       const int32_t* colidx_host, const float* values_host, int32_t nrhs,
                       spmvx_batch* out);
int spmvx_batch_destroy(spmvx_batch B);
/* Launch configuration: workgroup size (256, 512, 1024), lanes per row
 * (1, 2, 4, 8, 16; 0 = chosen per system from its mean row length) and rows
 * per team member: a system of n rows is solved by a team of
 * ceil(n / rows_per_member) workgroups (capped at 64 and at half the
 * workgroups resident on the device; 0 = one workgroup per system).  Set by
 * the autotuner (spmv_amd/autotune.py). */
int spmvx_batch_set_config(spmvx_batch B, int32_t block, int32_t lanes_per_row,
                           int32_t rows_per_member);
/*
 * Solves every system and right-hand side with CG from x = 0 until
 * ||r_k||^2 / ||b||^2 < rtol^2 or max_its, in one persistent kernel launch
 * (workgroups take team slots, largest systems first, from a device work
 * queue; team members synchronise through bounded waits), replayed from a
 * hipGraph captured on the first call with these arguments.  When the
 * recursive residual meets rtol, the true residual b - A x is formed in
 * double; a right-hand side whose true residual does not meet rtol restarts
 * CG from its x (r = p = b - A x) with a four times lower threshold for the
 * recursive residual, at most 4 times per solve, so the returned x meets
 * ||b - A x|| <= rtol ||b|| and not only the FP32 recursion.
 * its_host (may be NULL): nsys*nrhs iteration counts (counted on across
 * restarts).  The stream is synchronised; a timed-out team wait is reported
 * as SPMVX_ERR_STATE (results invalid).
 */
int spmvx_batch_solve(spmvx_batch B, const float* b_dev, float* x_dev, int32_t max_its,
                      double rtol, int32_t* its_host, spmvx_stream stream);
int spmvx_batch_info(spmvx_batch B, int32_t* nsys, int64_t* nrows, int64_t* nnz, int32_t* nrhs,
                     int32_t* block, int32_t* lanes_per_row, int32_t* rows_per_member,
                     int32_t* grid, int32_t* nslots);

/* ===================================================================== */
/* Extension 3: adjoint helpers (no reference counterpart)                */
/* ===================================================================== */
/* SDDMM on the pattern of A: out[k] = alpha * u[row(k)] * v[col(k)], for
 * F64 matrices (grad_A = -lambda x^T restricted to the pattern). */
int spmvx_sddmm(spmvx_matrix A, double alpha, const double* u_dev, const double* v_dev,
                double* out_dev, spmvx_stream stream);

/* ===================================================================== */
/* Utilities                                                              */
/* ===================================================================== */
/* dst[i] = src[i], n doubles: the known-size copy kernel used to calibrate
 * HBM byte counters (and the read+write ceiling).  No counterpart in the
 * reference (a measurement utility). */
int spmvx_copy_f64(double* dst_dev, const double* src_dev, int64_t n, spmvx_stream stream);
/* *sum_dev = the sum of n doubles (summation order unspecified), read once
 * with non-temporal 16-byte loads: the streaming-read ceiling of this GPU,
 * against which the SpMV kernels (~97% reads) are measured.  src_dev 16-byte
 * aligned.  No counterpart in the reference (a measurement utility). */
int spmvx_read_f64(const double* src_dev, int64_t n, double* sum_dev, spmvx_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* SPMV_AMD_H */
