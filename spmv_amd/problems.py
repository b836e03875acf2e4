"""Synthetic problems of the benchmark configurations (host, numpy).

* ``stencil``      5-point 2D / 7-point 3D / 27-point 3D Dirichlet stencils,
                   optionally with ghost z-planes owned by the next rank
                   (one rank's share of a z-slab decomposition)
* ``demo1d``       the 3-point matrix of demos/CreateA.cpp:52-74
* ``diffusion``    2D anisotropic 5-point diffusion from edge conductivities
* ``random_spd``   strictly diagonally dominant SPD with a random symmetric
                   pattern and power-law row lengths (batched solver)
* ``uniform`` / ``normal``  counter-based splitmix64 streams

Grid point (i, j, k) is row ``i + nx*(j + ny*k)``; columns are ascending
within every row (what Eigen's setFromTriplets produces).
"""
import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)


def rand_u64(seed, n, start=0):
    """splitmix64: value i of stream `seed` is mix(seed + (i+1)*gamma)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(start, start + n, dtype=np.uint64) + np.uint64(1)) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, n, lo=0.0, hi=1.0):
    u = (rand_u64(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return lo + (hi - lo) * u


def normal(seed, n):
    """Box-Muller on two uniform streams (seed, seed ^ 0x5DEECE66D)."""
    u1 = uniform(seed, n)
    u2 = uniform(seed ^ 0x5DEECE66D, n)
    u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


class CSR:
    """Host CSR (rowptr int32[nrows+1], colidx int32[nnz], val float64[nnz])."""

    def __init__(self, nrows, ncols, rowptr, colidx, val):
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.rowptr, self.colidx, self.val = rowptr, colidx, val

    @property
    def nnz(self):
        return int(self.rowptr[-1])


def _offsets(kind):
    offs = []
    for dk in (-1, 0, 1):
        for dj in (-1, 0, 1):
            for di in (-1, 0, 1):
                nz = (di != 0) + (dj != 0) + (dk != 0)
                if kind == 5 and (dk != 0 or nz > 1):
                    continue
                if kind == 7 and nz > 1:
                    continue
                offs.append((di, dj, dk))
    return offs  # ascending column order


def stencil(kind, nx, ny, nz=1, ghost_planes=0, diag=None, off=-1.0):
    """Box stencil with Dirichlet boundaries; `ghost_planes` extra z-planes
    above the top are ghost columns numbered after the owned ones."""
    if diag is None:
        diag = {5: 4.0, 7: 6.0, 27: 26.0}[kind]
    M = nx * ny * nz
    row = np.arange(M, dtype=np.int64)
    i = (row % nx).astype(np.int32)
    j = ((row // nx) % ny).astype(np.int32)
    k = (row // (nx * ny)).astype(np.int32)
    offs = _offsets(kind)
    S = len(offs)
    valid = np.empty((M, S), dtype=bool)
    cols = np.empty((M, S), dtype=np.int32)
    vals = np.empty(S, dtype=np.float64)
    for s, (di, dj, dk) in enumerate(offs):
        valid[:, s] = ((i + di >= 0) & (i + di < nx) & (j + dj >= 0) & (j + dj < ny)
                       & (k + dk >= 0) & (k + dk < nz + ghost_planes))
        cols[:, s] = (row + di + nx * dj + nx * ny * dk).astype(np.int32)
        vals[s] = diag if (di, dj, dk) == (0, 0, 0) else off
    counts = valid.sum(axis=1, dtype=np.int64)
    rowptr = np.zeros(M + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    colidx = cols[valid]
    del cols
    val = np.broadcast_to(vals, (M, S))[valid]
    return CSR(M, M + nx * ny * ghost_planes, rowptr.astype(np.int32), colidx,
               np.ascontiguousarray(val))


def demo1d(N, gamma=0.1):
    """demos/CreateA.cpp:52-74 on one rank."""
    rows, cols, vals = [], [], []
    r = np.arange(N)
    parts = []
    if N == 1:
        return CSR(1, 1, np.array([0, 1], np.int32), np.array([0], np.int32),
                   np.array([1.0 - gamma]))
    # interior rows: (g, 1-2g, g); first: (1-g, g); last: (g, 1-g)
    counts = np.full(N, 3, np.int64)
    counts[0] = counts[-1] = 2
    rowptr = np.zeros(N + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    colidx = np.empty(rowptr[-1], np.int32)
    val = np.empty(rowptr[-1], np.float64)
    colidx[0:2] = [0, 1]
    val[0:2] = [1.0 - gamma, gamma]
    mid = r[1:-1]
    base = rowptr[1:N - 1]
    for t, dc, v in ((0, -1, gamma), (1, 0, 1.0 - 2.0 * gamma), (2, 1, gamma)):
        colidx[base + t] = mid + dc
        val[base + t] = v
    colidx[-2:] = [N - 2, N - 1]
    val[-2:] = [gamma, 1.0 - gamma]
    del rows, cols, vals, parts
    return CSR(N, N, rowptr.astype(np.int32), colidx, val)


# ---------------------------------------------------------------- diffusion
def diffusion_num_edges(n):
    return 2 * n * (n - 1) + 4 * n


def diffusion_pattern(n):
    """Pattern of the n*n anisotropic 5-point operator and the edge incidence
    of every stored entry.

    Edges: horizontal (i,j)-(i+1,j) -> j*(n-1)+i; vertical (i,j)-(i,j+1) ->
    n(n-1) + j*n + i; boundary edges (Dirichlet) -> 2n(n-1) + side*n + t for
    side 0: j=0, 1: j=n-1, 2: i=0, 3: i=n-1.
    A = sum_interior theta_e (e_a-e_b)(e_a-e_b)^T + sum_boundary theta_e e_a e_a^T.

    Returns (CSR with zero values, P) where P is a host CSR of shape
    (nnz, n_edges) with A.values = P @ theta (+1 / -1 coefficients).
    """
    N = n * n
    nh = n * (n - 1)
    row = np.arange(N, dtype=np.int64)
    i, j = row % n, row // n
    e_s = np.where(j > 0, nh + (j - 1) * n + i, -1)
    e_w = np.where(i > 0, j * (n - 1) + (i - 1), -1)
    e_e = np.where(i < n - 1, j * (n - 1) + i, -1)
    e_n = np.where(j < n - 1, nh + j * n + i, -1)
    b0 = np.where(j == 0, 2 * nh + i, -1)
    b1 = np.where(j == n - 1, 2 * nh + n + i, -1)
    b2 = np.where(i == 0, 2 * nh + 2 * n + j, -1)
    b3 = np.where(i == n - 1, 2 * nh + 3 * n + j, -1)
    slot_cols = np.stack([row - n, row - 1, row, row + 1, row + n], axis=1)
    slot_edge = np.stack([e_s, e_w, np.zeros_like(row), e_e, e_n], axis=1)
    valid = np.ones((N, 5), dtype=bool)
    valid[:, 0] = e_s >= 0
    valid[:, 1] = e_w >= 0
    valid[:, 3] = e_e >= 0
    valid[:, 4] = e_n >= 0
    counts = valid.sum(1)
    rowptr = np.zeros(N + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    colidx = slot_cols[valid].astype(np.int32)
    nnz = int(rowptr[-1])
    # incidence: off-diagonal entry -> (edge, -1); diagonal -> all edges at a
    diag_edges = np.stack([e_s, e_w, e_e, e_n, b0, b1, b2, b3], axis=1)
    dvalid = diag_edges >= 0
    is_diag = np.zeros((N, 5), dtype=bool)
    is_diag[:, 2] = True
    entry_is_diag = is_diag[valid]
    entry_row = np.repeat(row, counts)
    entry_len = np.where(entry_is_diag, dvalid.sum(1)[entry_row], 1)
    prow = np.zeros(nnz + 1, np.int64)
    np.cumsum(entry_len, out=prow[1:])
    pcol = np.empty(prow[-1], np.int64)
    pval = np.empty(prow[-1], np.float64)
    off_idx = np.nonzero(~entry_is_diag)[0]
    pcol[prow[off_idx]] = slot_edge[valid][off_idx]
    pval[prow[off_idx]] = -1.0
    d_idx = np.nonzero(entry_is_diag)[0]
    de = diag_edges[entry_row[d_idx]]
    dv = dvalid[entry_row[d_idx]]
    # edges of the diagonal in ascending edge order
    de_sorted = np.sort(np.where(dv, de, np.iinfo(np.int64).max), axis=1)
    for t in range(8):
        m = t < entry_len[d_idx]
        pcol[prow[d_idx[m]] + t] = de_sorted[m, t]
        pval[prow[d_idx[m]] + t] = 1.0
    A = CSR(N, N, rowptr.astype(np.int32), colidx, np.zeros(nnz))
    P = CSR(nnz, diffusion_num_edges(n), prow.astype(np.int32), pcol.astype(np.int32), pval)
    return A, P


def csr_transpose(P):
    """Host transpose (rows ascending within each column)."""
    order = np.argsort(P.colidx, kind="stable")
    counts = np.bincount(P.colidx, minlength=P.ncols)
    rowptr = np.zeros(P.ncols + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    rows = np.repeat(np.arange(P.nrows), np.diff(P.rowptr.astype(np.int64)))
    return CSR(P.ncols, P.nrows, rowptr.astype(np.int32), rows[order].astype(np.int32),
               P.val[order].copy())


def diffusion(n, theta):
    """Assembled A(theta) (values = P @ theta, diagonal summed in ascending
    edge order)."""
    A, P = diffusion_pattern(n)
    theta = np.asarray(theta, np.float64)
    contrib = P.val * theta[P.colidx]
    start = P.rowptr[:-1].astype(np.int64)
    length = np.diff(P.rowptr.astype(np.int64))
    vals = np.zeros(A.nnz)
    for t in range(int(length.max()) if A.nnz else 0):  # sequential, slot by slot
        m = t < length
        vals[m] = vals[m] + contrib[start[m] + t]
    return CSR(A.nrows, A.ncols, A.rowptr, A.colidx, vals)


# ---------------------------------------------------------------- random SPD
def random_spd(n, seed, dmin=3, dmax=64, alpha=2.0):
    """Strictly diagonally dominant SPD: row i draws a target degree d_i from
    a power law P(d) ~ d^-alpha on [dmin, dmax]; it picks floor(d_i / 2)
    partners uniformly, every pair is mirrored (symmetric pattern), off
    diagonal values ~ U[-1, 0] shared by (i, j) and (j, i), and the diagonal
    is 1.1 * sum |off| (+1 for an empty row)."""
    u = uniform(seed, n)
    a1 = 1.0 - alpha
    d = ((dmin ** a1 + u * (dmax ** a1 - dmin ** a1)) ** (1.0 / a1)).astype(np.int64)
    d = np.clip(d, dmin, dmax)
    k = np.maximum(d // 2, 1)
    src = np.repeat(np.arange(n, dtype=np.int64), k)
    dst = (rand_u64(seed ^ 0xA5A5A5A5, len(src)) % np.uint64(max(n - 1, 1))).astype(np.int64)
    dst = np.where(dst >= src, dst + 1, dst) if n > 1 else src
    keep = dst != src
    src, dst = src[keep], dst[keep]
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    key = np.unique(lo * n + hi)
    lo, hi = key // n, key % n
    w = -uniform(seed ^ 0x3C3C3C3C, len(key))
    rows = np.concatenate([lo, hi])
    cols = np.concatenate([hi, lo])
    vals = np.concatenate([w, w])
    absrow = np.bincount(rows, weights=-vals, minlength=n)
    diag = 1.1 * absrow
    diag[diag == 0.0] = 1.0
    rows = np.concatenate([rows, np.arange(n)])
    cols = np.concatenate([cols, np.arange(n)])
    vals = np.concatenate([vals, diag])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=n), out=rowptr[1:])
    return CSR(n, n, rowptr.astype(np.int32), cols.astype(np.int32), vals)


# ---------------------------------------------------------------- halo plan
def slab_decomposition(nx, ny, nz, ghost_planes=1):
    """The owned block of rank 0 in a z-slab decomposition of an
    nx*ny*(2*nz) grid into two ranks: ghosts are the first `ghost_planes`
    planes of rank 1, and rank 1 ghosts the top `ghost_planes` planes of
    rank 0.  Returns kwargs for spmv_amd.L2GMap."""
    M = nx * ny * nz
    plane = nx * ny
    ghosts = np.arange(M, M + plane * ghost_planes, dtype=np.int64)
    remote = {1: np.arange(M - plane * ghost_planes, M, dtype=np.int64)}
    return dict(local_size=M, ghosts=ghosts, local_sizes=[M, M], rank=0, remote_ghosts=remote)
