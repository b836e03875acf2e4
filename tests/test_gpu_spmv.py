"""GPU parity: SpMV, transpose SpMV, L2GMap update / reverse_update and SDDMM
through the C ABI against the CPU oracle.  FP64 results must be bit-identical
(the reference's Eigen row-major product sums every row sequentially with
separately rounded products and sums; so does the oracle)."""
import numpy as np
import pytest
import torch

import oracle as O
import spmv_amd as S
from spmv_amd import problems as P

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=DEV, dtype=dtype)


def _ocsr(A):
    return O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val)


def long_row_matrix(n=6000, seed=4):
    """Rows of very different lengths: a dense first row (> the 2048-entry
    LDS chunk), empty rows, and short random rows."""
    rng = np.random.default_rng(seed)
    rows = [np.arange(n, dtype=np.int32)]
    for i in range(1, n):
        if i % 7 == 0:
            rows.append(np.zeros(0, np.int32))
        elif i % 1001 == 0:
            rows.append(np.sort(rng.choice(n, 3000, replace=False)).astype(np.int32))
        else:
            rows.append(np.sort(rng.choice(n, rng.integers(1, 40), replace=False)).astype(np.int32))
    rowptr = np.zeros(n + 1, np.int64)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    colidx = np.concatenate(rows)
    val = P.uniform(seed, len(colidx), -1, 1)
    return P.CSR(n, n, rowptr.astype(np.int32), colidx, val)


def matrices():
    yield "demo1d", P.demo1d(64)
    yield "poisson5", P.stencil(5, 33, 31)
    yield "poisson7", P.stencil(7, 13, 12, 11)
    yield "stencil27_ghost", P.stencil(27, 11, 12, 13, ghost_planes=1)
    yield "diffusion", P.diffusion(40, np.exp(0.5 * P.normal(3, P.diffusion_num_edges(40))))
    yield "random_spd", P.random_spd(5000, 11)
    yield "long_rows", long_row_matrix()
    yield "one_row", P.CSR(1, 3, np.array([0, 3], np.int32), np.array([0, 1, 2], np.int32),
                           np.array([1.0, -2.0, 3.0]))


MATS = list(matrices())


@pytest.mark.parametrize("name,A", MATS, ids=[m[0] for m in MATS])
def test_spmv_f64_bitexact(name, A):
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    x = P.uniform(7, A.ncols, -1, 1)
    y = M.mult(_dev(x)).cpu().numpy()
    assert np.array_equal(y, O.spmv(_ocsr(A), x))


@pytest.mark.parametrize("name,A", MATS, ids=[m[0] for m in MATS])
def test_spmv_transpose_bitexact(name, A):
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    x = P.uniform(8, A.nrows, -1, 1)
    yt = M.transpmult(_dev(x)).cpu().numpy()
    assert np.array_equal(yt, O.spmv_t(_ocsr(A), x))


def _rect_matrix(nrows=3000, ncols=4500, seed=31):
    """Non-symmetric rectangular matrix with empty rows, empty columns (the
    last 500 are never referenced) and rows of 0-60 random entries."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(nrows):
        k = 0 if i % 11 == 0 else int(rng.integers(1, 60))
        rows.append(np.sort(rng.choice(ncols - 500, k, replace=False)).astype(np.int32))
    rowptr = np.zeros(nrows + 1, np.int64)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    colidx = np.concatenate(rows)
    return P.CSR(nrows, ncols, rowptr.astype(np.int32), colidx,
                 P.uniform(seed, len(colidx), -1, 1))


@pytest.mark.parametrize("dtype", ["f64", "c64"])
def test_transpose_rectangular_and_value_updates_on_side_stream(dtype):
    """The transposed copy is built on the device (stable sort by column) on
    the caller's stream: bit-exact against the oracle's transposed product on
    a non-symmetric rectangular matrix with empty rows and columns, and after
    value updates made on a side stream (non-blocking, as torch's are) and
    followed by transposed products on that same stream: first before the
    copy exists, then with the copy kept and its values re-gathered."""
    A = _rect_matrix()
    oc = _ocsr(A)
    x = P.uniform(41, A.nrows, -1, 1)
    side = torch.cuda.Stream()
    if dtype == "f64":
        M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
        assert np.array_equal(M.transpmult(_dev(x)).cpu().numpy(), O.spmv_t(oc, x))
        for seed in (42, 43):
            v2 = P.uniform(seed, A.nnz, -3, 3)
            v2d = _dev(v2)
            xd = _dev(x)
            torch.cuda.synchronize()
            with torch.cuda.stream(side):
                M.set_values(v2d)
                yt = M.transpmult(xd)
            side.synchronize()
            B = P.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, v2)
            assert np.array_equal(yt.cpu().numpy(), O.spmv_t(_ocsr(B), x))
            assert np.array_equal(M.mult(_dev(P.uniform(seed, A.ncols, -1, 1))).cpu().numpy(),
                                  O.spmv(_ocsr(B), P.uniform(seed, A.ncols, -1, 1)))
    else:
        v = A.val + 1j * P.uniform(44, A.nnz, -1, 1)
        M = S.Matrix(A.rowptr, A.colidx, v, ncols=A.ncols, dtype="c64")
        xc = x + 1j * P.uniform(45, A.nrows, -1, 1)
        assert np.array_equal(M.transpmult(_cdev(xc)).cpu().numpy(), O.spmv_t_complex(oc, v, xc))
        v2 = P.uniform(46, A.nnz, -1, 1) - 1j * P.uniform(47, A.nnz, -1, 1)
        v2d = _cdev(v2).view(torch.float64)
        xd = _cdev(xc)
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            M.set_values(v2d)
            yt = M.transpmult(xd)
        side.synchronize()
        assert np.array_equal(yt.cpu().numpy(), O.spmv_t_complex(oc, v2, xc))


def _bf16_round(v):
    """double -> bf16 (round to nearest even) -> float32, on the host."""
    f = v.astype(np.float32)
    # exact double rounding via round-to-odd in float
    back = f.astype(np.float64)
    bits = f.view(np.uint32).copy()
    inexact = back != v
    toward = np.abs(back) > np.abs(v)
    bits = np.where(inexact & toward, bits - 1, bits)
    bits = np.where(inexact, bits | 1, bits).astype(np.uint32)
    bits = (bits + 0x7FFF + ((bits >> 16) & 1)) >> 16
    return (bits.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("dtype", ["f32", "bf16"])
@pytest.mark.parametrize("name,A", MATS[:7], ids=[m[0] for m in MATS[:7]])
def test_spmv_low_precision(dtype, name, A):
    M64 = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    x = P.uniform(9, A.ncols, -1, 1).astype(np.float32)
    v32 = A.val.astype(np.float32) if dtype == "f32" else _bf16_round(A.val)
    want = O.spmv_f32(_ocsr(A), v32, x)
    for M in (S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols, dtype=dtype), M64.like(dtype)):
        y = M.mult(_dev(x, torch.float32)).cpu().numpy()
        assert np.array_equal(y, want)


def test_spmv_full_size_27pt_ghost_plane():
    """One rank's 160^3 share of the 27-point stencil with its ghost plane
    (109,443,836 stored entries), bit-exact against the oracle."""
    A = P.stencil(27, 160, 160, 160, ghost_planes=1)
    assert A.nnz == 109_443_836
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    x = P.uniform(21, A.ncols, -1, 1)
    y = M.mult(_dev(x)).cpu().numpy()
    assert np.array_equal(y, O.spmv(_ocsr(A), x, threads=8))


@pytest.mark.parametrize("variant", [-1, 5, S.SLICED_VARIANT])
def test_spmv_timed_matches_product(variant):
    """spmvx_spmv_timed: the same y as the product, and two events recorded by
    the kernel itself (elapsed time positive and within the bracketing
    events' span); without entries they bracket the zero fill."""
    A = P.stencil(27, 24, 20, 22, ghost_planes=1)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    M.set_kernel(variant)
    x = _dev(P.uniform(8, A.ncols, -1, 1))
    y0 = M.mult(x)
    y1 = torch.full_like(y0, 7.0)
    e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    s = torch.cuda.current_stream()
    e[0].record(s)
    M.mult_timed(x, y1, e[1], e[2])
    e[3].record(s)
    torch.cuda.synchronize()
    assert torch.equal(y0, y1)
    k, outer = e[1].elapsed_time(e[2]), e[0].elapsed_time(e[3])
    assert 0 < k <= outer
    Z = S.Matrix(np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0), ncols=2)
    z = torch.full((3,), 5.0, dtype=torch.float64, device=DEV)
    Z.mult_timed(torch.ones(2, dtype=torch.float64, device=DEV), z, e[1], e[2])
    torch.cuda.synchronize()
    assert torch.all(z == 0) and e[1].elapsed_time(e[2]) >= 0


def test_spmv_empty_matrix():
    M = S.Matrix(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), ncols=0)
    y = M.mult(torch.zeros(0, dtype=torch.float64, device=DEV))
    assert y.numel() == 0


# ---------------------------------------------------------------- L2GMap
def test_update_pack_unpack():
    kw = P.slab_decomposition(9, 8, 7)
    m = S.L2GMap(**kw)
    n = m.local_size() + m.num_ghosts()
    v = P.uniform(1, n, -1, 1)
    halo = P.uniform(2, m.num_ghosts(), -5, 5)
    m.set_halo(halo)
    vd = _dev(v)
    send = m.update(vd, send_out=True)
    vo = v.copy()
    send_o = O.l2g_update(m.local_size(), m.num_ghosts(), m.plan()["indexbuf"], vo, halo)
    assert np.array_equal(vd.cpu().numpy(), vo)
    assert np.array_equal(send, send_o)
    # halo reset to zeros
    m.set_halo(None)
    m.update(vd)
    assert np.all(vd.cpu().numpy()[m.local_size():] == 0.0)


def test_update_float32():
    kw = P.slab_decomposition(5, 4, 6)
    m = S.L2GMap(**kw)
    n = m.local_size() + m.num_ghosts()
    v = P.uniform(3, n, -1, 1).astype(np.float32)
    halo = P.uniform(4, m.num_ghosts(), -1, 1).astype(np.float32)
    m.set_halo(halo, dtype=S.F32)
    vd = _dev(v, torch.float32)
    send = m.update(vd, send_out=True)
    assert np.array_equal(send, v[m.plan()["indexbuf"]])
    assert np.array_equal(vd.cpu().numpy()[m.local_size():], halo)


def test_reverse_update_repeated_indices():
    # rank 1 of 3 in the "scatter" decomposition: indexbuf repeats indices
    import json, os
    gold = [json.load(open(os.path.join(os.path.dirname(__file__), "golden", "l2gmap",
                                        f"scatter_3_{r}.json"))) for r in range(3)]
    r = 1
    g = gold[r]
    lo, hi = g["ranges"][r], g["ranges"][r + 1]
    remote = {q: [x for x in gold[q]["ghosts"] if lo <= x < hi] for q in range(3) if q != r}
    m = S.L2GMap(g["local_size"], g["ghosts"], local_sizes=[x["local_size"] for x in gold], rank=r,
                 remote_ghosts=remote)
    idx = m.plan()["indexbuf"]
    assert len(set(idx)) < len(idx)
    n = m.local_size() + m.num_ghosts()
    v = P.uniform(5, n, -1, 1)
    inc = P.uniform(6, len(idx), -1, 1)
    vd = _dev(v)
    out = m.reverse_update(vd, inc)
    vo = v.copy()
    out_o = O.l2g_reverse_update(m.local_size(), m.num_ghosts(), idx, vo, inc)
    assert np.array_equal(out, out_o)
    assert np.array_equal(vd.cpu().numpy(), vo)


# ---------------------------------------------------------------- SDDMM
def test_sddmm():
    A = P.stencil(5, 17, 19)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    u = P.uniform(1, A.nrows, -1, 1)
    v = P.uniform(2, A.ncols, -1, 1)
    g = S.sddmm(M, _dev(u), _dev(v), alpha=-1.0).cpu().numpy()
    rows = np.repeat(np.arange(A.nrows), np.diff(A.rowptr))
    assert np.array_equal(g, (-u[rows]) * v[A.colidx])


@pytest.mark.parametrize("name,A", [MATS[3], MATS[5], MATS[6]], ids=["stencil27_ghost", "random_spd", "long_rows"])
def test_all_kernel_variants_bitexact(name, A):
    want = O.spmv(_ocsr(A), P.uniform(7, A.ncols, -1, 1))
    x = _dev(P.uniform(7, A.ncols, -1, 1))
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    Mb = M.like("bf16")
    xf = x.float()
    ref_b = None
    for v in range(S.lib().spmvx_num_kernels()):
        if v == S.SLICED_VARIANT and name != "stencil27_ghost":
            continue  # no fit: covered by test_sliced_layout_rejects
        M.set_kernel(v)
        assert M.kernel()[0] == v and M.kernel()[1]
        assert np.array_equal(M.mult(x).cpu().numpy(), want), v
        Mb.set_kernel(v)
        yb = Mb.mult(xf).cpu().numpy()
        if ref_b is None:
            ref_b = yb
        assert np.array_equal(yb, ref_b), v


# ---------------------------------------------------------------- rows longer than a block
def very_long_rows(n=40000, seed=12):
    """Every row holds column 0 (so column 0 of A, row 0 of A^T, is dense);
    row 0 is dense; rows 1, 7 and 4000 hold 30000, 12000 and 8500 entries
    (past every variant's 2048-8192-entry block); the rest 1-30 entries."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if i == 0:
            c = np.arange(n)
        else:
            k = {1: 30000, 7: 12000, 4000: 8500}.get(i, int(rng.integers(1, 30)))
            c = np.union1d([0], rng.choice(np.arange(1, n), k, replace=False))
        rows.append(c.astype(np.int32))
    rowptr = np.zeros(n + 1, np.int64)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    colidx = np.concatenate(rows)
    val = P.uniform(seed, len(colidx), -1, 1)
    return P.CSR(n, n, rowptr.astype(np.int32), colidx, val)


def _exact_rows(A, val, x, dtype):
    """Per row: the exact sum (math.fsum) of the products rounded in `dtype`,
    and sum |a_ij x_j|."""
    import math
    prod = (val.astype(dtype) * x.astype(dtype)[A.colidx]).astype(np.float64)
    ex = np.array([math.fsum(prod[A.rowptr[i]:A.rowptr[i + 1]]) for i in range(A.nrows)])
    mag = np.add.reduceat(np.abs(prod), A.rowptr[:-1]) * (np.diff(A.rowptr) > 0)
    return ex, mag


def _check_long_rows(y, ex, mag, lens, u):
    """Rows past every block (> 8192 entries) are summed by a workgroup of at
    least 256 threads: each thread adds at most ceil(len/256) products, then
    a tree of depth <= 10 (plus the final round) combines them.  Recursive
    summation of m terms errs by at most m*u*sum|terms| (to first order)."""
    long = lens > 8192
    assert long.any()
    m = -(-lens[long] // 256) + 12
    assert np.all(np.abs(y[long] - ex[long]) <= 1.01 * m * u * mag[long])


LONG = very_long_rows()


def test_long_rows_ordered_all_variants_bitexact():
    A = LONG
    x = P.uniform(3, A.ncols, -1, 1)
    want = O.spmv(_ocsr(A), x)
    want_t = O.spmv_t(_ocsr(A), x)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    for v in range(S.SLICED_VARIANT):
        M.set_kernel(v)
        assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), want), v
    assert np.array_equal(M.transpmult(_dev(x)).cpu().numpy(), want_t)


def test_long_rows_parallel_all_variants():
    """Parallel long rows: deterministic; rows that fit a block unchanged
    (bit-exact); long rows within the summation bound of the exact sum."""
    A = LONG
    lens = np.diff(A.rowptr)
    x = P.uniform(3, A.ncols, -1, 1)
    want = O.spmv(_ocsr(A), x)
    ex, mag = _exact_rows(A, A.val, x, np.float64)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    M.set_long_rows(True)
    u = 2.0 ** -53
    for v in range(S.SLICED_VARIANT):
        M.set_kernel(v)
        y = M.mult(_dev(x)).cpu().numpy()
        assert np.array_equal(y, M.mult(_dev(x)).cpu().numpy()), v   # deterministic
        short = lens <= 2040
        assert np.array_equal(y[short], want[short]), v
        _check_long_rows(y, ex, mag, lens, u)
        # the mode is in effect: some long row is rounded differently
        assert not np.array_equal(y[lens > 8192], want[lens > 8192]), v
    # transposed product: row 0 of A^T holds all 40000 rows
    import scipy.sparse as sp
    T = sp.csr_matrix((A.val, A.colidx, A.rowptr), shape=(A.nrows, A.ncols)).T.tocsr()
    At = P.CSR(A.ncols, A.nrows, T.indptr.astype(np.int32), T.indices.astype(np.int32), T.data)
    ext, magt = _exact_rows(At, At.val, x, np.float64)
    yt = M.transpmult(_dev(x)).cpu().numpy()
    lt = np.diff(At.rowptr)
    assert np.array_equal(yt[lt <= 2040], O.spmv_t(_ocsr(A), x)[lt <= 2040])
    _check_long_rows(yt, ext, magt, lt, u)
    # back to the ordered sum: bit-exact again
    M.set_long_rows(False)
    M.set_kernel(-1)
    assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), want)


def test_long_rows_parallel_f32_and_restrictions():
    A = LONG
    lens = np.diff(A.rowptr)
    x = P.uniform(4, A.ncols, -1, 1).astype(np.float32)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    M.set_long_rows(True)
    Mf = M.like("f32")      # inherits the mode
    y = Mf.mult(_dev(x, torch.float32)).cpu().numpy().astype(np.float64)
    ex, mag = _exact_rows(A, A.val.astype(np.float32), x, np.float32)
    _check_long_rows(y, ex, mag, lens, 2.0 ** -24)
    Mc = S.Matrix(A.rowptr, A.colidx, A.val.astype(np.complex128), ncols=A.ncols, dtype="c64")
    with pytest.raises(S.SpmvError, match="real values"):
        Mc.set_long_rows(True)
    with pytest.raises(S.SpmvError, match="0 or 1"):
        S._check(S.lib().spmvx_matrix_set_long_rows(M._h, 2))


def test_long_rows_cg_both_modes():
    """CG on an SPD arrow matrix whose first row and column are dense (20000
    entries): both modes converge to the direct solution."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = 20000
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 0, n - 1) / n
    off = sp.coo_matrix((np.concatenate([u, u]),
                         (np.concatenate([np.zeros(n - 1, int), np.arange(1, n)]),
                          np.concatenate([np.arange(1, n), np.zeros(n - 1, int)]))), shape=(n, n))
    T = sp.diags([-0.3 * np.ones(n - 1), -0.3 * np.ones(n - 1)], [-1, 1])
    Asp = (off + T).tocsr()
    Asp = (Asp + sp.diags(1.1 * np.asarray(abs(Asp).sum(axis=1)).ravel() + 1.0)).tocsr()
    Asp.sort_indices()
    b = P.uniform(6, n, -1, 1)
    xd = spla.spsolve(Asp.tocsc(), b)
    its = []
    for par in (False, True):
        M = S.Matrix(Asp.indptr.astype(np.int32), Asp.indices.astype(np.int32), Asp.data, ncols=n)
        M.set_long_rows(par)
        x, k = S.cg(M, _dev(b), 500, 1e-12)
        x = x.cpu().numpy()[:n]
        assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
        its.append(k)
    assert abs(its[0] - its[1]) <= 2


# ---------------------------------------------------------------- 16-bit column indices
SHORT_VARIANTS = [8, 9, 10, 11]


def _mixed_window_matrix(n=150000, seed=51):
    """A band of width 41 around the diagonal, rows near n/3 that also reach
    three far columns (their row blocks need four 16384-column windows, the
    others one), empty rows, and one row longer than every staging chunk (a
    block of its own, summed from the 32-bit indices)."""
    rows = []
    for i in range(n):
        lo, hi = max(0, i - 20), min(n, i + 21)
        cols = set(range(lo, hi))
        if n // 3 <= i < n // 3 + 300:
            cols |= {(i + 40000 * k) % n for k in range(1, 4)}
        if i % 97 == 5:
            cols = set()
        if i == n // 2:
            cols = set(range(0, n, 8))  # 18750 entries
        rows.append(np.array(sorted(cols), np.int32))
    rowptr = np.zeros(n + 1, np.int64)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    colidx = np.concatenate(rows)
    return P.CSR(n, n, rowptr.astype(np.int32), colidx, P.uniform(seed, len(colidx), -1, 1))


def test_short_index_variants_bitexact():
    """The 16-bit column-index kernels decode the same columns in the same
    order: FP64 products and transposed products bit-identical to the oracle,
    FP32 / BF16 bit-identical to their 32-bit-index kernels, on a matrix whose
    row blocks need one to four windows, with empty rows and a row longer
    than the staging chunk."""
    A = _mixed_window_matrix()
    x = P.uniform(52, A.ncols, -1, 1)
    want = O.spmv(_ocsr(A), x)
    xt = P.uniform(53, A.nrows, -1, 1)
    want_t = O.spmv_t(_ocsr(A), xt)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    xf = _dev(x, torch.float32)
    ref = {dt: None for dt in ("f32", "bf16")}
    for dt in ref:
        Ml = M.like(dt)
        Ml.set_kernel(3 if dt == "f32" else 4)
        ref[dt] = Ml.mult(xf).cpu().numpy()
    for v in SHORT_VARIANTS:
        M.set_kernel(v)
        assert "16-bit" in M.kernel_name() and "windows" in M.kernel_name()
        assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), want), v
        assert np.array_equal(M.transpmult(_dev(xt)).cpu().numpy(), want_t), v
        for dt in ref:
            Ml = M.like(dt)
            Ml.set_kernel(v)
            assert np.array_equal(Ml.mult(xf).cpu().numpy(), ref[dt]), (v, dt)


def test_short_index_defaults():
    """Stencil and band matrices get a 16-bit-index kernel by default (FP64,
    FP32 and BF16 values): one base per row block when a block's columns span
    less than 65536, else up to four 16384-column windows per block (the
    7-point operator on 192 x 192 x 20, whose plane neighbours are 36864
    columns away: three windows); a matrix with uniformly random columns over
    a wide range keeps 32-bit indices, as do complex values."""
    for A, kind in ((P.stencil(27, 20, 20, 20, ghost_planes=1), "one base"),
                    (P.stencil(7, 192, 192, 20), "windows"),
                    (P.stencil(5, 200, 200), "one base")):
        M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
        assert "16-bit" in M.kernel_name() and kind in M.kernel_name()
        for dt in ("f32", "bf16"):
            assert "16-bit" in M.like(dt).kernel_name()
        x = P.uniform(54, A.ncols, -1, 1)
        assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), O.spmv(_ocsr(A), x))
        xt = P.uniform(55, A.nrows, -1, 1)
        assert np.array_equal(M.transpmult(_dev(xt)).cpu().numpy(), O.spmv_t(_ocsr(A), xt))
    # uniformly random columns over 200000: no block fits four 16384-column
    # windows
    R = P.random_spd(200000, 3)
    MR = S.Matrix(R.rowptr, R.colidx, R.val, ncols=R.ncols)
    assert "16-bit" not in MR.kernel_name()
    with pytest.raises(S.SpmvError, match="16-bit"):
        MR.set_kernel(SHORT_VARIANTS[0])
    C = P.stencil(7, 8, 8, 8)
    assert "16-bit" not in S.Matrix(C.rowptr, C.colidx, C.val.astype(complex), ncols=C.ncols,
                                    dtype="c64").kernel_name()


# ---------------------------------------------------------------- sliced row-pattern layout
SLICED_FIT = ["demo1d", "poisson5", "poisson7", "stencil27_ghost", "diffusion", "one_row"]


@pytest.mark.parametrize("name,A", [m for m in MATS if m[0] in SLICED_FIT], ids=SLICED_FIT)
def test_sliced_layout_bitexact(name, A):
    """The sliced layout sums every row in CSR order: FP64 bit-identical to
    the oracle, FP32/BF16 bit-identical to the CSR kernels, transposed
    products unchanged (they use the CSR arrays)."""
    x = P.uniform(17, A.ncols, -1, 1)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    M.set_kernel(S.SLICED_VARIANT)
    assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), O.spmv(_ocsr(A), x))
    xt = P.uniform(18, A.nrows, -1, 1)
    assert np.array_equal(M.transpmult(_dev(xt)).cpu().numpy(), O.spmv_t(_ocsr(A), xt))
    xf = _dev(x, torch.float32)
    for dt in ("f32", "bf16"):
        Ml = M.like(dt)
        ref = Ml.mult(xf).cpu().numpy()
        Ml.set_kernel(S.SLICED_VARIANT)
        assert np.array_equal(Ml.mult(xf).cpu().numpy(), ref), dt


def band_matrix(n, lo, hi, seed):
    """Rows i hold columns i-lo .. i+hi (clipped at the edges): one interior
    pattern of width lo+hi+1 and lo+hi truncated ones."""
    rows = [np.arange(max(0, i - lo), min(n, i + hi + 1), dtype=np.int32) for i in range(n)]
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    colidx = np.concatenate(rows)
    return P.CSR(n, n, rowptr, colidx, P.uniform(seed, len(colidx), -1, 1))


SLICED_WIDTHS = [(1, 1), (2, 2), (4, 4), (7, 7), (8, 7), (8, 8), (14, 14), (20, 20)]
# slots per step the launcher picks for each width above (fewest idle slots,
# then fewest steps, among the compiled steps; one value per load, no
# padding): FP64 4/7/8/10/14, FP32 4/7/8/14, BF16 4/7
SLICED_STEPS = {"f64": [4, 7, 10, 8, 8, 10, 10, 14], "f32": [4, 7, 4, 8, 8, 4, 8, 14],
                "bf16": [4, 7, 4, 4, 4, 4, 4, 7]}
_steps_seen = {dt: set() for dt in SLICED_STEPS}


def _slots_per_step(M):
    import re
    return int(re.search(r"(\d+) slots per step", M.kernel_name()).group(1))


@pytest.mark.parametrize("case", range(len(SLICED_WIDTHS)),
                         ids=["w3", "w5", "w9", "w15", "w16", "w17", "w29", "w41"])
def test_sliced_layout_pattern_widths(case):
    """Band matrices of pattern widths 3-41 on both sides of the sliced
    kernels' per-step slot counts: the step the launcher picks (reported by
    the kernel description) is the expected one, FP64 is bit-identical to the
    oracle and FP32/BF16 to the CSR kernels; over the cases every compiled
    step of every value type is used."""
    lo, hi = SLICED_WIDTHS[case]
    A = band_matrix(3001, lo, hi, seed=lo * 100 + hi)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    M.set_kernel(S.SLICED_VARIANT)
    assert M.sliced_info()["wmax"] == lo + hi + 1
    assert _slots_per_step(M) == SLICED_STEPS["f64"][case]
    _steps_seen["f64"].add(_slots_per_step(M))
    x = P.uniform(23, A.ncols, -1, 1)
    assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), O.spmv(_ocsr(A), x))
    xf = _dev(x, torch.float32)
    for dt in ("f32", "bf16"):
        Ml = M.like(dt)
        ref = Ml.mult(xf).cpu().numpy()
        Ml.set_kernel(S.SLICED_VARIANT)
        assert _slots_per_step(Ml) == SLICED_STEPS[dt][case], dt
        _steps_seen[dt].add(_slots_per_step(Ml))
        assert np.array_equal(Ml.mult(xf).cpu().numpy(), ref), dt


def test_sliced_layout_every_step_used():
    """The width cases above cover every compiled step (runs after them)."""
    if not all(_steps_seen.values()):
        pytest.skip("runs after the width cases")
    assert _steps_seen == {"f64": {4, 7, 8, 10, 14}, "f32": {4, 7, 8, 14}, "bf16": {4, 7}}


def test_sliced_layout_info():
    # 7-point stencil on a 13 x 12 x 11 grid: 3 boundary classes per axis
    A = P.stencil(7, 13, 12, 11)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    info = M.sliced_info()
    assert info["npatterns"] == 27 and info["wmax"] == 7
    lens = np.diff(A.rowptr)
    nsl = -(-A.nrows // 64)
    w = [lens[64 * i:64 * (i + 1)].max() for i in range(nsl)]
    assert info["slots"] == 64 * sum(w)  # FP64: one entry per lane and load
    assert M.like("f32").sliced_info()["slots"] == 64 * sum(w)  # FP32 too
    assert M.like("bf16").sliced_info()["slots"] == 64 * sum(w)  # and BF16


@pytest.mark.parametrize("name,A,reason", [
    ("random_spd", MATS[5][1], "pattern"),
    ("long_rows", MATS[6][1], "longer"),
], ids=["random_spd", "long_rows"])
def test_sliced_layout_rejects(name, A, reason):
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    with pytest.raises(S.SpmvError, match=reason):
        M.set_kernel(S.SLICED_VARIANT)
    with pytest.raises(S.SpmvError, match=reason):
        M.sliced_info()
    x = P.uniform(19, A.ncols, -1, 1)
    assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), O.spmv(_ocsr(A), x))


def test_sliced_layout_set_values():
    A = MATS[4][1]  # diffusion
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    Mb = M.like("bf16")
    M.set_kernel(S.SLICED_VARIANT)
    Mb.set_kernel(S.SLICED_VARIANT)
    v2 = P.uniform(20, A.nnz, -2, 2)
    M.set_values(_dev(v2))
    Mb.set_values(_dev(v2))
    torch.cuda.synchronize()
    x = P.uniform(21, A.ncols, -1, 1)
    B = P.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, v2)
    assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), O.spmv(_ocsr(B), x))
    Mc = S.Matrix(A.rowptr, A.colidx, v2, ncols=A.ncols, dtype="bf16")
    xf = _dev(x, torch.float32)
    assert np.array_equal(Mb.mult(xf).cpu().numpy(), Mc.mult(xf).cpu().numpy())


# ---------------------------------------------------------------- complex values
def _cdev(a, dtype=torch.complex128):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=DEV, dtype=dtype)


@pytest.mark.parametrize("name,A", [m for m in MATS if m[0] in
                                    ("poisson7", "stencil27_ghost", "random_spd", "long_rows", "one_row")],
                         ids=["poisson7", "stencil27_ghost", "random_spd", "long_rows", "one_row"])
def test_complex_spmv_bitexact(name, A):
    """Matrix<std::complex<double>> / <float> products: bit-identical to the
    oracle's sequential complex row sums (which equal scipy's), every kernel
    variant that takes complex values, and the transposed product."""
    v = A.val + 1j * P.uniform(13, A.nnz, -1, 1)
    x = P.uniform(14, A.ncols, -1, 1) + 1j * P.uniform(15, A.ncols, -1, 1)
    oc = _ocsr(A)
    want = O.spmv_complex(oc, v, x)
    M = S.Matrix(A.rowptr, A.colidx, v, ncols=A.ncols, dtype="c64")
    assert M.vector_dtype() == torch.complex128
    xd = _cdev(x)
    for variant in range(5):
        M.set_kernel(variant)
        assert np.array_equal(M.mult(xd).cpu().numpy(), want), variant
    xt = P.uniform(16, A.nrows, -1, 1) + 1j * P.uniform(17, A.nrows, -1, 1)
    assert np.array_equal(M.transpmult(_cdev(xt)).cpu().numpy(), O.spmv_t_complex(oc, v, xt))
    M32 = S.Matrix(A.rowptr, A.colidx, v, ncols=A.ncols, dtype="c32")
    x32 = x.astype(np.complex64)
    want32 = O.spmv_complex(oc, v.astype(np.complex64), x32)
    assert np.array_equal(M32.mult(_cdev(x32, torch.complex64)).cpu().numpy(), want32)


def test_complex_restrictions_and_values():
    A = MATS[2][1]  # poisson7
    v = A.val + 0.5j
    M = S.Matrix(A.rowptr, A.colidx, v, ncols=A.ncols, dtype="c64")
    for bad in (5, 11, S.SLICED_VARIANT):
        with pytest.raises(S.SpmvError, match="complex|real values"):
            M.set_kernel(bad)
    with pytest.raises(S.SpmvError, match="real values"):
        M.sliced_info()
    with pytest.raises(S.SpmvError):
        S.cg(M, _cdev(np.ones(A.nrows)), 10, 1e-8)
    # new values from the device ((re, im) pairs)
    v2 = P.uniform(18, A.nnz, -1, 1) - 2j * P.uniform(19, A.nnz, -1, 1)
    M.set_values(_cdev(v2))
    x = P.uniform(20, A.ncols, -1, 1) + 1j * P.uniform(21, A.ncols, -1, 1)
    assert np.array_equal(M.mult(_cdev(x)).cpu().numpy(), O.spmv_complex(_ocsr(A), v2, x))


@pytest.mark.parametrize("tdt", [torch.complex128, torch.complex64])
def test_complex_update_and_reverse(tdt):
    """L2GMap<std::complex<T>>::update / reverse_update (L2GMap.cpp:202-213):
    pure data movement and ordered scatter-add on complex entries."""
    import json, os
    gold = [json.load(open(os.path.join(os.path.dirname(__file__), "golden", "l2gmap",
                                        f"scatter_3_{r}.json"))) for r in range(3)]
    rank = 1
    g = gold[rank]
    lo, hi = g["ranges"][rank], g["ranges"][rank + 1]
    remote = {q: [v for v in gold[q]["ghosts"] if lo <= v < hi] for q in range(3) if q != rank}
    m = S.L2GMap(g["local_size"], g["ghosts"], local_sizes=[x["local_size"] for x in gold],
                 rank=rank, remote_ghosts=remote)
    n, ng = m.local_size(), m.num_ghosts()
    npdt = np.complex128 if tdt == torch.complex128 else np.complex64
    v = (P.uniform(22, n + ng, -1, 1) + 1j * P.uniform(23, n + ng, -1, 1)).astype(npdt)
    halo = (P.uniform(24, ng, -1, 1) - 1j * P.uniform(25, ng, -1, 1)).astype(npdt)
    m.set_halo(halo, dtype=S.C64 if tdt == torch.complex128 else S.C32)
    vd = _cdev(v, tdt)
    send = m.update(vd, send_out=True)
    idx = m.plan()["indexbuf"]
    want = v.copy()
    want[n:] = halo
    assert np.array_equal(vd.cpu().numpy(), want)
    assert np.array_equal(send, v[np.asarray(idx, np.int64)])
    inc = (P.uniform(26, len(idx), -1, 1) + 1j * P.uniform(27, len(idx), -1, 1)).astype(npdt)
    out = m.reverse_update(vd, inc)
    assert np.array_equal(out, halo)
    acc = want.copy()
    for j, t in enumerate(idx):  # sequential order of L2GMap.cpp:168-169
        acc[t] = acc[t] + inc[j]
    assert np.array_equal(vd.cpu().numpy()[:n], acc[:n])


def test_create_like_copies_from_two_threads():
    """Matrices made by create_like share their pattern; two host threads
    that select different kernel variants (each building its own row blocks
    on the shared pattern on first use) give the sequential results."""
    import threading
    A = P.stencil(27, 24, 24, 24)
    M64 = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    x = P.uniform(31, A.ncols, -1, 1)
    want64 = O.spmv(_ocsr(A), x)
    xf = x.astype(np.float32)
    want32 = O.spmv_f32(_ocsr(A), A.val.astype(np.float32), xf)
    errors = []

    def work(M, xv, want, variants):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for v in variants:
                    M.set_kernel(v)
                    y = M.mult(xv, stream=s)
                    s.synchronize()
                    if not np.array_equal(y.cpu().numpy(), want):
                        errors.append(("variant", v))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(repr(e))

    M32 = M64.like("f32")
    xd, xfd = _dev(x), _dev(xf, torch.float32)
    torch.cuda.synchronize()  # inputs on the device before the side streams read them
    t1 = threading.Thread(target=work, args=(M64, xd, want64, [0, 5, 8, 11, 12, 7, 9]))
    t2 = threading.Thread(target=work, args=(M32, xfd, want32, [12, 3, 7, 10, 1, 6, 2]))
    t1.start()
    t2.start()
    t1.join()
    t2.join()
    assert not errors, errors


def test_products_reject_overlapping_vectors():
    """The product reads x while it writes y; overlapping ranges are an
    argument error, adjacent ones are fine."""
    A = P.stencil(5, 8, 8)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    x = _dev(P.uniform(41, A.ncols, -1, 1))
    with pytest.raises(S.SpmvError, match="overlap"):
        M.mult(x, out=x)
    with pytest.raises(S.SpmvError, match="overlap"):
        M.transpmult(x, out=x)
    buf = torch.zeros(A.ncols + A.nrows, dtype=torch.float64, device=DEV)
    buf[:A.ncols] = x
    y = M.mult(buf[:A.ncols], out=buf[A.ncols:])
    assert np.array_equal(y.cpu().numpy(), O.spmv(_ocsr(A), x.cpu().numpy()))


def boundary_matrix(seed):
    """Random CSR whose row lengths sit on and around every kernel's block
    limits (2040/2048, 4088/4096, 8184/8192 entries), mixed with empty rows
    and runs of short rows, so blocks end exactly at, just before and just
    after their limits."""
    rng = np.random.default_rng(seed)
    ncols = int(rng.integers(9000, 20000))
    specials = [0, 1, 7, 2039, 2040, 2041, 2047, 2048, 2049, 4087, 4088, 4089, 4095, 4096, 4097,
                8183, 8184, 8185, 8191, 8192, 8193]
    # short rows (10% empty), then every special length twice at random places
    lens = list(np.where(rng.random(3000) < 0.1, 0, rng.integers(1, 60, 3000)))
    for ln in specials + specials:
        lens.insert(int(rng.integers(0, len(lens) + 1)), ln)
    lens = np.minimum(np.array(lens, np.int64), ncols)
    rowptr = np.zeros(len(lens) + 1, np.int64)
    rowptr[1:] = np.cumsum(lens)
    colidx = np.concatenate([np.sort(rng.choice(ncols, l, replace=False)) if l else
                             np.zeros(0, np.int64) for l in lens]).astype(np.int32)
    val = P.uniform(seed + 50, len(colidx), -1, 1)
    return P.CSR(len(lens), ncols, rowptr.astype(np.int32), colidx, val)


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5, 6])
def test_block_boundary_rows_every_variant(seed):
    A = boundary_matrix(seed)
    x = P.uniform(seed + 60, A.ncols, -1, 1)
    want = O.spmv(_ocsr(A), x)
    want_t = O.spmv_t(_ocsr(A), P.uniform(seed + 61, A.nrows, -1, 1))
    xf = x.astype(np.float32)
    want32 = O.spmv_f32(_ocsr(A), A.val.astype(np.float32), xf)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    Mf = M.like("f32")
    for v in range(S.SLICED_VARIANT):
        M.set_kernel(v)
        assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), want), (seed, v)
        Mf.set_kernel(v)
        assert np.array_equal(Mf.mult(_dev(xf, torch.float32)).cpu().numpy(), want32), (seed, v)
    assert np.array_equal(M.transpmult(_dev(P.uniform(seed + 61, A.nrows, -1, 1))).cpu().numpy(),
                          want_t)


def test_products_of_matrices_without_entries():
    """No stored entries: y = 0, including a matrix with rows but no columns
    (x empty) and the transposed product of a matrix without rows; no product
    kernel runs (its branch-free gathers would need x to hold an entry)."""
    M = S.Matrix(np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0), ncols=0)
    y = M.mult(torch.zeros(0, dtype=torch.float64, device=DEV))
    assert y.shape == (5,) and torch.all(y == 0)
    Mf = M.like("f32")
    assert torch.all(Mf.mult(torch.zeros(0, dtype=torch.float32, device=DEV)) == 0)
    Mt = S.Matrix(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), ncols=4)
    yt = Mt.transpmult(torch.zeros(0, dtype=torch.float64, device=DEV))
    assert yt.shape == (4,) and torch.all(yt == 0)
    Ms = S.Matrix(np.zeros(7, np.int32), np.zeros(0, np.int32), np.zeros(0), ncols=6)
    out = torch.full((6,), 3.0, dtype=torch.float64, device=DEV)
    Ms.mult(torch.ones(6, dtype=torch.float64, device=DEV), out=out)
    assert torch.all(out == 0)


def test_spmv_autotuner_cache(tmp_path, monkeypatch):
    """tune_spmv times every variant the matrix fits, applies the fastest
    and caches it; a second call uses the cache; a cached choice that does
    not apply to the matrix at hand (the sliced layout for a matrix it does
    not fit, or a variant whose kernel differs from the stored one) is tuned
    again instead of failing.  Results are bit-exact throughout."""
    from spmv_amd import autotune as T
    monkeypatch.setenv("SPMV_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    A = P.stencil(27, 20, 20, 20)
    x = P.uniform(51, A.ncols, -1, 1)
    want = O.spmv(_ocsr(A), x)
    M = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
    best, table = T.tune_spmv(M, _dev(x), reps=2)
    assert table is not None and len(table) == S.lib().spmvx_num_kernels()  # the stencil fits all
    assert M.kernel() == (best["variant"], best["kernel"])
    assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), want)
    best2, table2 = T.tune_spmv(M, _dev(x), reps=2)
    assert table2 is None and best2 == best
    # planted entries for a matrix of the same signature
    R = P.random_spd(3000, 8)
    MR = S.Matrix(R.rowptr, R.colidx, R.val, ncols=R.ncols)
    key = T.spmv_signature(MR)
    xr = P.uniform(52, R.ncols, -1, 1)
    for planted in ({"variant": S.SLICED_VARIANT, "kernel": "sliced", "ms": 0.0},
                    {"variant": 0, "kernel": "not this kernel", "ms": 0.0}):
        T._store(key, planted)
        bestr, tabler = T.tune_spmv(MR, _dev(xr), reps=1)
        assert tabler is not None
        assert all(r["variant"] != S.SLICED_VARIANT for r in tabler)  # refused, skipped
        assert MR.kernel() == (bestr["variant"], bestr["kernel"])
        assert np.array_equal(MR.mult(_dev(xr)).cpu().numpy(), O.spmv(_ocsr(R), xr))
