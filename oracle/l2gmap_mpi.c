/*
 * Real-MPI restatement of the L2GMap constructor and update, used ONLY to
 * write golden fixtures (tests/golden/make_l2gmap_golden.py runs it under
 * mpirun -n 1..4 in a container with MPICH).  TEST INFRASTRUCTURE ONLY.
 *
 * Follows spmv/L2GMap.cpp:15-122 (constructor: Allgather of sizes, owner
 * lookup, Alltoall symmetrisation, dist-graph comm, Neighbor_alltoall of
 * counts, Neighbor_alltoallv of ghost-local indices, send_offset += local
 * size) and spmv/L2GMap.cpp:126-150 (update: gather indexbuf then
 * Neighbor_alltoallv with the count roles reversed), and :152-170
 * (reverse_update: ghost region to owners, scatter-add).
 *
 * Every rank writes <outdir>/<case>_<size>_<rank>.json with its plan and the
 * result of one update and one reverse_update on v[i] = global index
 * (owned) / -1 (ghost).
 *
 * Problem cases (rows partitioned with owner_ranges, demos/CreateA.cpp:12-29):
 *   demo1d : N=37, 3-point rows of demos/CreateA.cpp:52-74
 *   grid2d : 7x7 5-point stencil
 *   slab27 : 5x4x(3*size+1) 27-point stencil
 *   scatter: N=41, row i -> cols {i, (7i+3)%N, (i*i+5)%N} (asymmetric
 *            neighbour sets, exercises the symmetrisation of :57-78)
 */
#include <assert.h>
#include <mpi.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void owner_ranges(int size, int64_t N, int64_t* ranges)
{
  const int64_t n = N / size, r = N % size;
  for (int rank = 0; rank < size + 1; ++rank)
    ranges[rank] = (rank < r) ? rank * (n + 1) : rank * n + r;
}

/* Column list of global row `row` for each case; returns count. */
static int case_cols(const char* name, int64_t row, int64_t N, int64_t* cols)
{
  int c = 0;
  if (!strcmp(name, "demo1d"))
  {
    if (row > 0) cols[c++] = row - 1;
    cols[c++] = row;
    if (row < N - 1) cols[c++] = row + 1;
  }
  else if (!strcmp(name, "grid2d"))
  {
    const int64_t n = 7, i = row % n, j = row / n;
    if (j > 0) cols[c++] = row - n;
    if (i > 0) cols[c++] = row - 1;
    cols[c++] = row;
    if (i < n - 1) cols[c++] = row + 1;
    if (j < n - 1) cols[c++] = row + n;
  }
  else if (!strcmp(name, "slab27"))
  {
    const int64_t nx = 5, ny = 4, nz = N / (nx * ny);
    const int64_t i = row % nx, j = (row / nx) % ny, k = row / (nx * ny);
    for (int dk = -1; dk <= 1; ++dk)
      for (int dj = -1; dj <= 1; ++dj)
        for (int di = -1; di <= 1; ++di)
        {
          const int64_t ii = i + di, jj = j + dj, kk = k + dk;
          if (ii < 0 || ii >= nx || jj < 0 || jj >= ny || kk < 0 || kk >= nz)
            continue;
          cols[c++] = ii + nx * (jj + ny * kk);
        }
  }
  else /* scatter */
  {
    cols[c++] = row;
    cols[c++] = (7 * row + 3) % N;
    cols[c++] = (row * row + 5) % N;
  }
  return c;
}

static int64_t case_size(const char* name, int size)
{
  if (!strcmp(name, "demo1d")) return 37;
  if (!strcmp(name, "grid2d")) return 49;
  if (!strcmp(name, "slab27")) return 5 * 4 * (3 * (int64_t)size + 1);
  return 41;
}

static int cmp64(const void* a, const void* b)
{
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

static void dump_i32(FILE* f, const char* key, const int32_t* a, int n, int last)
{
  fprintf(f, "  \"%s\": [", key);
  for (int i = 0; i < n; ++i)
    fprintf(f, "%s%d", i ? ", " : "", a[i]);
  fprintf(f, "]%s\n", last ? "" : ",");
}
static void dump_i64(FILE* f, const char* key, const int64_t* a, int n, int last)
{
  fprintf(f, "  \"%s\": [", key);
  for (int i = 0; i < n; ++i)
    fprintf(f, "%s%lld", i ? ", " : "", (long long)a[i]);
  fprintf(f, "]%s\n", last ? "" : ",");
}
static void dump_f64(FILE* f, const char* key, const double* a, int n, int last)
{
  fprintf(f, "  \"%s\": [", key);
  for (int i = 0; i < n; ++i)
    fprintf(f, "%s%.17g", i ? ", " : "", a[i]);
  fprintf(f, "]%s\n", last ? "" : ",");
}

int main(int argc, char** argv)
{
  MPI_Init(&argc, &argv);
  if (argc != 3)
  {
    fprintf(stderr, "usage: l2gmap_mpi <case> <outdir>\n");
    MPI_Abort(MPI_COMM_WORLD, 1);
  }
  const char* name = argv[1];
  int mpi_size, mpi_rank;
  MPI_Comm comm = MPI_COMM_WORLD;
  MPI_Comm_size(comm, &mpi_size);
  MPI_Comm_rank(comm, &mpi_rank);

  /* ---- the rank's rows and its ghost columns (CreateA.cpp:77-87) ---- */
  const int64_t N = case_size(name, mpi_size);
  int64_t* rows = malloc(sizeof(int64_t) * (mpi_size + 1));
  owner_ranges(mpi_size, N, rows);
  const int64_t local_size = rows[mpi_rank + 1] - rows[mpi_rank];
  int64_t* ghosts = malloc(sizeof(int64_t) * (size_t)(local_size * 27 + 1));
  int nghost = 0;
  for (int64_t row = rows[mpi_rank]; row < rows[mpi_rank + 1]; ++row)
  {
    int64_t cols[27];
    int c = case_cols(name, row, N, cols);
    for (int t = 0; t < c; ++t)
      if (cols[t] < rows[mpi_rank] || cols[t] >= rows[mpi_rank + 1])
        ghosts[nghost++] = cols[t];
  }
  qsort(ghosts, (size_t)nghost, sizeof(int64_t), cmp64);
  { /* unique */
    int u = 0;
    for (int i = 0; i < nghost; ++i)
      if (u == 0 || ghosts[u - 1] != ghosts[i])
        ghosts[u++] = ghosts[i];
    nghost = u;
  }

  /* ---- constructor, L2GMap.cpp:19-121 ---- */
  int64_t* ranges = calloc((size_t)mpi_size + 1, sizeof(int64_t));
  int64_t ls = local_size;
  MPI_Allgather(&ls, 1, MPI_INT64_T, ranges + 1, 1, MPI_INT64_T, comm);
  for (int i = 0; i < mpi_size; ++i)
    ranges[i + 1] += ranges[i];
  const int64_t r0 = ranges[mpi_rank], r1 = ranges[mpi_rank + 1];

  int32_t* ghost_count = calloc((size_t)mpi_size, sizeof(int32_t));
  int32_t* ghost_local = malloc(sizeof(int32_t) * (size_t)(nghost + 1));
  for (int i = 0; i < nghost; ++i)
  {
    const int64_t idx = ghosts[i];
    assert(!(idx >= r0 && idx < r1));
    int p = 0; /* upper_bound(ranges, idx) - 1 */
    while (p + 1 <= mpi_size && ranges[p + 1] <= idx)
      ++p;
    ++ghost_count[p];
    ghost_local[i] = (int32_t)(idx - ranges[p]);
  }

  int32_t* remote_count = calloc((size_t)mpi_size, sizeof(int32_t));
  MPI_Alltoall(ghost_count, 1, MPI_INT, remote_count, 1, MPI_INT, comm);

  int* neighbours = malloc(sizeof(int) * (size_t)(mpi_size + 1));
  int32_t* send_count = malloc(sizeof(int32_t) * (size_t)(mpi_size + 1));
  int nn = 0;
  for (int i = 0; i < mpi_size; ++i)
  {
    if (ghost_count[i] > 0)
    {
      neighbours[nn] = i;
      send_count[nn++] = ghost_count[i];
    }
    else if (remote_count[i] > 0)
    {
      neighbours[nn] = i;
      send_count[nn++] = 0;
    }
  }
  MPI_Comm ncomm;
  MPI_Dist_graph_create_adjacent(comm, nn, neighbours, MPI_UNWEIGHTED, nn,
                                 neighbours, MPI_UNWEIGHTED, MPI_INFO_NULL, 0,
                                 &ncomm);
  int32_t* recv_count = calloc((size_t)nn + 1, sizeof(int32_t));
  int nsc = nn;
  if (nn == 0)
  {
    send_count[0] = 0;
    recv_count[0] = 0;
    nsc = 1;
  }
  MPI_Neighbor_alltoall(send_count, 1, MPI_INT, recv_count, 1, MPI_INT, ncomm);
  int32_t* send_offset = malloc(sizeof(int32_t) * (size_t)(nsc + 1));
  int32_t* recv_offset = malloc(sizeof(int32_t) * (size_t)(nsc + 1));
  send_offset[0] = recv_offset[0] = 0;
  for (int i = 0; i < nsc; ++i)
  {
    send_offset[i + 1] = send_offset[i] + send_count[i];
    recv_offset[i + 1] = recv_offset[i] + recv_count[i];
  }
  const int count = recv_offset[nsc];
  int32_t* indexbuf = malloc(sizeof(int32_t) * (size_t)(count + 1));
  MPI_Neighbor_alltoallv(ghost_local, send_count, send_offset, MPI_INT32_T,
                         indexbuf, recv_count, recv_offset, MPI_INT32_T,
                         ncomm);
  for (int i = 0; i < nsc + 1; ++i)
    send_offset[i] += (int32_t)local_size;

  /* ---- update, L2GMap.cpp:126-150 ---- */
  const int64_t nv = local_size + nghost;
  double* v = malloc(sizeof(double) * (size_t)(nv + 1));
  for (int64_t i = 0; i < local_size; ++i)
    v[i] = (double)(r0 + i);
  for (int64_t i = local_size; i < nv; ++i)
    v[i] = -1.0;
  double* databuf = malloc(sizeof(double) * (size_t)(count + 1));
  for (int i = 0; i < count; ++i)
    databuf[i] = v[indexbuf[i]];
  MPI_Neighbor_alltoallv(databuf, recv_count, recv_offset, MPI_DOUBLE, v,
                         send_count, send_offset, MPI_DOUBLE, ncomm);

  /* ---- reverse_update, L2GMap.cpp:152-170, on w: owned 0, ghost g -> g+1 */
  double* w = malloc(sizeof(double) * (size_t)(nv + 1));
  for (int64_t i = 0; i < local_size; ++i)
    w[i] = 0.0;
  for (int64_t i = local_size; i < nv; ++i)
    w[i] = (double)(1000 * (mpi_rank + 1) + (i - local_size));
  MPI_Neighbor_alltoallv(w, send_count, send_offset, MPI_DOUBLE, databuf,
                         recv_count, recv_offset, MPI_DOUBLE, ncomm);
  for (int i = 0; i < count; ++i)
    w[indexbuf[i]] += databuf[i];

  char path[4096];
  snprintf(path, sizeof(path), "%s/%s_%d_%d.json", argv[2], name, mpi_size,
           mpi_rank);
  FILE* f = fopen(path, "w");
  fprintf(f, "{\n  \"case\": \"%s\",\n  \"size\": %d,\n  \"rank\": %d,\n",
          name, mpi_size, mpi_rank);
  fprintf(f, "  \"local_size\": %lld,\n", (long long)local_size);
  dump_i64(f, "ranges", ranges, mpi_size + 1, 0);
  dump_i64(f, "ghosts", ghosts, nghost, 0);
  dump_i32(f, "neighbours", neighbours, nn, 0);
  dump_i32(f, "send_count", send_count, nsc, 0);
  dump_i32(f, "recv_count", recv_count, nsc, 0);
  dump_i32(f, "send_offset", send_offset, nsc + 1, 0);
  dump_i32(f, "recv_offset", recv_offset, nsc + 1, 0);
  dump_i32(f, "indexbuf", indexbuf, count, 0);
  dump_f64(f, "update_v", v, (int)nv, 0);
  dump_f64(f, "reverse_w", w, (int)nv, 1);
  fprintf(f, "}\n");
  fclose(f);

  MPI_Comm_free(&ncomm);
  MPI_Finalize();
  return 0;
}
