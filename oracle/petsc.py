"""PETSc binary files: a numpy restatement of spmv/read_petsc.cpp, plus a
writer of the same layout.  TEST INFRASTRUCTURE ONLY (tests/ and the file
fixtures they write); the library's reader is spmv_amd/csrc/petsc.cpp.

Layout (big-endian, 32-bit indices; what PETSc writes with -ksp_view_mat /
-ksp_view_rhs binary):
    matrix: int32 1211216, rows, cols, nnz | int32 row lengths[rows] |
            int32 column indices[nnz] | float64 values[nnz]
    vector: int32 1211214, rows | float64 values[rows]
"""
import numpy as np

MAT_CLASSID = 1211216  # read_petsc.cpp:65
VEC_CLASSID = 1211214  # read_petsc.cpp:220


def owner_ranges(size, N):
    """read_petsc.cpp:13-30: N items in `size` contiguous ranges, the first
    N % size of them one longer."""
    n, r = divmod(N, size)
    return [q * (n + 1) if q < r else q * n + r for q in range(size + 1)]


def write_matrix(path, nrows, ncols, rowptr, colidx, values):
    """Rows in the given order; column indices as given (PETSc writes them
    sorted, but the reader sorts each row itself)."""
    rowptr = np.asarray(rowptr, dtype=np.int64)
    with open(path, "wb") as f:
        f.write(np.array([MAT_CLASSID, nrows, ncols, rowptr[-1]], dtype=">i4").tobytes())
        f.write(np.diff(rowptr).astype(">i4").tobytes())
        f.write(np.asarray(colidx).astype(">i4").tobytes())
        f.write(np.asarray(values, dtype=np.float64).astype(">f8").tobytes())


def write_vector(path, values):
    values = np.asarray(values, dtype=np.float64)
    with open(path, "wb") as f:
        f.write(np.array([VEC_CLASSID, values.size], dtype=">i4").tobytes())
        f.write(values.astype(">f8").tobytes())


def read_matrix(path, nranks=1, rank=0):
    """read_petsc.cpp:32-191 for one rank: returns dict(rowptr, colidx,
    values, ghosts, nrows_local, ncols_local, nrows, ncols, row_ranges,
    col_ranges, remote) where colidx is the local numbering (owned columns
    first, ghosts after in ascending global order, :115-133) and every row is
    sorted by local column (Eigen's insert, :152-160); remote[q] lists the
    columns of this rank that rank q ghosts."""
    with open(path, "rb") as f:
        data = f.read()
    head = np.frombuffer(data[:16], dtype=">i4").astype(np.int64)
    if head[0] != MAT_CLASSID:
        raise ValueError("Bad signature in PETSc Matrix file")
    nrows, ncols, nnz = (int(v) for v in head[1:])
    off = 16
    lens = np.frombuffer(data[off:off + 4 * nrows], dtype=">i4").astype(np.int64)
    off += 4 * nrows
    cols = np.frombuffer(data[off:off + 4 * nnz], dtype=">i4").astype(np.int64)
    off += 4 * nnz
    vals = np.frombuffer(data[off:off + 8 * nnz], dtype=">f8").astype(np.float64)
    start = np.concatenate([[0], np.cumsum(lens)])
    rr, cr = owner_ranges(nranks, nrows), owner_ranges(nranks, ncols)
    r0, r1, c0, c1 = rr[rank], rr[rank + 1], cr[rank], cr[rank + 1]
    k0, k1 = start[r0], start[r1]
    lc, lv = cols[k0:k1], vals[k0:k1]
    ghosts = np.unique(lc[(lc < c0) | (lc >= c1)])
    local = np.where((lc >= c0) & (lc < c1), lc - c0, (c1 - c0) + np.searchsorted(ghosts, lc))
    rowptr = (start[r0:r1 + 1] - k0).astype(np.int64)
    colidx = np.empty_like(local)
    values = np.empty_like(lv)
    for i in range(r1 - r0):
        a, e = rowptr[i], rowptr[i + 1]
        order = np.argsort(local[a:e], kind="stable")
        colidx[a:e] = local[a:e][order]
        values[a:e] = lv[a:e][order]
    remote = []
    for q in range(nranks):
        if q == rank:
            remote.append(np.zeros(0, np.int64))
            continue
        qc = cols[start[rr[q]]:start[rr[q + 1]]]
        remote.append(np.unique(qc[(qc >= c0) & (qc < c1)]))
    return {"rowptr": rowptr, "colidx": colidx, "values": values, "ghosts": ghosts,
            "nrows_local": r1 - r0, "ncols_local": c1 - c0, "nrows": nrows, "ncols": ncols,
            "row_ranges": rr, "col_ranges": cr, "remote": remote}


def read_vector(path, nranks=1, rank=0):
    """read_petsc.cpp:193-260: this rank's contiguous part."""
    with open(path, "rb") as f:
        data = f.read()
    head = np.frombuffer(data[:8], dtype=">i4").astype(np.int64)
    if head[0] != VEC_CLASSID:
        raise ValueError("Bad signature in PETSc Vector file")
    n = int(head[1])
    r = owner_ranges(nranks, n)
    return np.frombuffer(data[8 + 8 * r[rank]:8 + 8 * r[rank + 1]], dtype=">f8").astype(np.float64)
