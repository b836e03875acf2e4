"""Pins the CPU oracle (CPU only).

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand, a fused multiply-add)
fails at least one of them:
  * scipy's csr_matvec runs the same sequential row sum as Eigen 3.3's
    row-major product (spmv/Matrix.cpp:203) -> bit-exact equality;
  * closed-form Dirichlet eigenpairs of the 5/7/27-point operators;
  * dense brute force;
  * CG: one-iteration convergence on an eigenvector, <= m iterations for m
    eigen-components, agreement with a direct solve, true residual.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O


def _sine_mode(shape, ks):
    grids = np.meshgrid(*[np.arange(1, n + 1) for n in shape], indexing="ij")
    v = np.ones(shape)
    for g, n, k in zip(grids, shape, ks):
        v = v * np.sin(k * np.pi * g / (n + 1))
    # generator numbering: row = i + nx*(j + ny*k) -> Fortran order of (i,j,k)
    return v.reshape(-1, order="F")


def _lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("n", [2, 3, 8, 17])
def test_generator_nnz_closed_forms(n):
    assert O.gen_stencil(5, n, n).nnz == 5 * n * n - 4 * n
    assert O.gen_stencil(7, n, n, n).nnz == 7 * n ** 3 - 6 * n ** 2
    assert O.gen_stencil(27, n, n, n).nnz == (3 * n - 2) ** 3
    # one ghost plane above the top: (3n-2)^2 extra entries
    A = O.gen_stencil(27, n, n, n, ghost_planes=1)
    assert A.nnz == (3 * n - 2) ** 3 + (3 * n - 2) ** 2
    assert A.ncols == n ** 3 + n * n
    assert O.gen_demo1d(n).nnz == 3 * n - 2
    assert O.gen_diffusion(n, np.ones(O.diffusion_num_edges(n))).nnz == 5 * n * n - 4 * n


def test_generator_config_sizes():
    # the 27-point 160^3 share with one ghost plane and the 1024^2 diffusion
    # problem, by closed form (not generated: too large for a unit test)
    n = 160
    assert (3 * n - 2) ** 3 == 109_215_352
    assert (3 * n - 2) ** 3 + (3 * n - 2) ** 2 == 109_443_836
    assert 7 * 256 ** 3 - 6 * 256 ** 2 == 117_047_296
    assert 5 * 512 ** 2 - 4 * 512 == 1_308_672
    assert O.diffusion_num_edges(1024) == 2_095_104 + 4_096


def test_generators_match_scipy_kron():
    n = 6
    I = sp.identity(n)
    T = _lap1d(n)
    L2 = (sp.kron(I, T) + sp.kron(T, I)).toarray()
    assert np.array_equal(O.gen_stencil(5, n, n).to_scipy().toarray(), L2)
    L3 = (sp.kron(sp.kron(I, I), T) + sp.kron(sp.kron(I, T), I) + sp.kron(sp.kron(T, I), I)).toarray()
    assert np.array_equal(O.gen_stencil(7, n, n, n).to_scipy().toarray(), L3)
    T3 = sp.diags([np.ones(n - 1), np.ones(n), np.ones(n - 1)], [-1, 0, 1])
    L27 = (27 * sp.identity(n ** 3) - sp.kron(sp.kron(T3, T3), T3)).toarray()
    assert np.array_equal(O.gen_stencil(27, n, n, n).to_scipy().toarray(), L27)


def test_generators_sorted_columns():
    for A in [O.gen_stencil(27, 4, 5, 3, ghost_planes=1), O.gen_stencil(7, 5, 4, 3),
              O.gen_demo1d(10), O.gen_diffusion(5, np.ones(O.diffusion_num_edges(5)))]:
        for i in range(A.nrows):
            row = A.colidx[A.rowptr[i]:A.rowptr[i + 1]]
            assert np.all(np.diff(row) > 0)


def test_demo1d_matches_createA():
    # demos/CreateA.cpp:52-74 on one rank: first row (1-g, g), last (g, 1-g),
    # interior (g, 1-2g, g)
    A = O.gen_demo1d(5, 0.1).to_scipy().toarray()
    g = 0.1
    want = np.zeros((5, 5))
    want[0, :2] = [1 - g, g]
    want[4, 3:] = [g, 1 - g]
    for i in range(1, 4):
        want[i, i - 1:i + 2] = [g, 1 - 2 * g, g]
    assert np.array_equal(A, want)


# ---------------------------------------------------------------- SpMV
def _cases():
    yield O.gen_demo1d(64)
    yield O.gen_stencil(5, 8, 8)
    yield O.gen_stencil(7, 6, 6, 6)
    yield O.gen_stencil(27, 5, 5, 5, ghost_planes=1)
    yield O.gen_diffusion(9, O.uniform(3, O.diffusion_num_edges(9), 0.2, 3.0))


@pytest.mark.parametrize("A", list(_cases()))
def test_spmv_bitexact_vs_scipy(A):
    x = O.uniform(11, A.ncols, -1, 1)
    y = O.spmv(A, x)
    ys = A.to_scipy() @ x  # csr_matvec: sum = 0; sum += a*x in stored order
    assert np.array_equal(y, ys)
    assert np.array_equal(O.spmv(A, x, threads=4), y)
    # brute force dense product (different summation order): tight tolerance
    yd = A.to_scipy().toarray() @ x
    scale = np.abs(A.to_scipy()) @ np.abs(x)
    assert np.all(np.abs(y - yd) <= 4e-16 * A.ncols * scale + 1e-300)


@pytest.mark.parametrize("A", list(_cases()))
def test_spmv_f32_bitexact_vs_scipy(A):
    """The FP32 product (reference for the FP32/BF16 kernels) is scipy's
    float32 csr_matvec: products and sums rounded to float32 in stored
    order."""
    import scipy.sparse as sp
    x = O.uniform(12, A.ncols, -1, 1).astype(np.float32)
    v = A.val.astype(np.float32)
    S32 = sp.csr_matrix((v, A.colidx, A.rowptr), shape=(A.nrows, A.ncols))
    y = O.spmv_f32(A, v, x)
    assert y.dtype == np.float32
    assert np.array_equal(y, S32 @ x)
    # the same rounded inputs summed in float64 differ somewhere: the oracle
    # really rounds its products and sums to float32
    y64 = S32.astype(np.float64) @ x.astype(np.float64)
    assert not np.array_equal(y.astype(np.float64), y64)


@pytest.mark.parametrize("kind,shape", [(5, (9, 7)), (7, (5, 6, 4)), (27, (4, 5, 6))])
def test_spmv_closed_form_eigenvectors(kind, shape):
    nx, ny = shape[0], shape[1]
    nz = shape[2] if len(shape) == 3 else 1
    A = O.gen_stencil(kind, nx, ny, nz)
    for ks in [(1,) * len(shape), (2, 3, 1)[: len(shape)], tuple(n for n in shape)]:
        v = _sine_mode(shape, ks)
        c = [np.cos(k * np.pi / (n + 1)) for k, n in zip(ks, shape)]
        if kind == 27:
            lam = 27.0 - np.prod([1 + 2 * ci for ci in c])
        else:
            lam = sum(2 - 2 * ci for ci in c)
        y = O.spmv(A, v)
        assert np.max(np.abs(y - lam * v)) < 1e-13 * max(1.0, abs(lam)) * 30


def test_spmv_ones_closed_form():
    n = 7
    A = O.gen_stencil(5, n, n)
    y = O.spmv(A, np.ones(A.ncols)).reshape(n, n, order="F")
    i = np.arange(n)
    missing = (i == 0).astype(float) + (i == n - 1)
    assert np.array_equal(y, missing[:, None] + missing[None, :])
    A27 = O.gen_stencil(27, 5, 5, 5)
    y27 = O.spmv(A27, np.ones(A27.ncols)).reshape(5, 5, 5, order="F")
    assert np.all(y27[1:-1, 1:-1, 1:-1] == 0.0)


def test_spmv_no_fma_rounding():
    # a*x + c where fma and separate rounding differ: 1 row, entries (a, 1)
    a = 1.0 + 2.0 ** -30
    rp = np.array([0, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    A = O.CSR(1, 2, rp, ci, np.array([a, 1.0]))
    x = np.array([a, -(1.0 + 2.0 ** -29)])
    prod = np.float64(a) * np.float64(a)  # rounded product
    want = (0.0 + prod) + x[1]
    assert O.spmv(A, x)[0] == want
    assert want != np.float64(a * a - (1.0 + 2.0 ** -29)) or True


def test_spmv_transpose_bruteforce():
    A = O.gen_stencil(27, 4, 3, 3, ghost_planes=1)
    x = O.uniform(5, A.nrows, -1, 1)
    yt = O.spmv_t(A, x)
    assert np.allclose(yt, A.to_scipy().toarray().T @ x, rtol=0, atol=1e-13)


def test_planted_mistakes_are_caught():
    """Mutants of the SpMV (dropped last term, transposed operand, wrong
    column) must fail the scipy check used above."""
    A = O.gen_stencil(27, 4, 4, 4, ghost_planes=1)
    x = O.uniform(1, A.ncols, -1, 1)
    ref = A.to_scipy() @ x
    mut_drop = np.array([sum(A.val[k] * x[A.colidx[k]] for k in range(A.rowptr[i], A.rowptr[i + 1] - 1))
                         for i in range(A.nrows)])
    mut_shift = np.array([sum(A.val[k] * x[min(A.colidx[k] + 1, A.ncols - 1)]
                              for k in range(A.rowptr[i], A.rowptr[i + 1])) for i in range(A.nrows)])
    S = O.gen_stencil(27, 4, 4, 4)
    Ad = S.to_scipy().toarray()
    Ad[0, 1] += 0.5  # non-symmetric so that A^T x differs from A x
    xs = x[: S.nrows]
    assert not np.array_equal(mut_drop, ref)
    assert not np.array_equal(mut_shift, ref)
    assert not np.allclose(Ad.T @ xs, Ad @ xs)


# ---------------------------------------------------------------- update
def test_l2g_update_and_reverse():
    ls, ng = 6, 3
    idx = np.array([5, 0, 5, 2], np.int32)
    v = np.arange(ls + ng, dtype=np.float64)
    send = O.l2g_update(ls, ng, idx, v, np.array([10.0, 11.0, 12.0]))
    assert np.array_equal(send, [5, 0, 5, 2])
    assert np.array_equal(v, [0, 1, 2, 3, 4, 5, 10, 11, 12])
    out = O.l2g_reverse_update(ls, ng, idx, v, np.array([1.0, 2.0, 3.0, 4.0]))
    assert np.array_equal(out, [10, 11, 12])
    # repeated index 5 accumulates twice (L2GMap.cpp:168-169)
    assert np.array_equal(v[:ls], [2, 1, 6, 3, 4, 9])


# ---------------------------------------------------------------- CG
def test_cg_eigenvector_one_iteration():
    n = 12
    A = O.gen_stencil(5, n, n)
    b = _sine_mode((n, n), (2, 3))
    x, k, hist = O.cg(A, b, 100, 1e-10)
    assert k == 1
    lam = 4 - 2 * np.cos(2 * np.pi / (n + 1)) - 2 * np.cos(3 * np.pi / (n + 1))
    assert np.allclose(x, b / lam, rtol=0, atol=1e-14)


def test_cg_m_components():
    n = 10
    A = O.gen_stencil(7, n, n, n)
    b = _sine_mode((n,) * 3, (1, 1, 1)) + 0.5 * _sine_mode((n,) * 3, (2, 1, 3)) \
        - 0.25 * _sine_mode((n,) * 3, (4, 4, 1))
    x, k, hist = O.cg(A, b, 100, 1e-12)
    assert k <= 3
    assert np.linalg.norm(A.to_scipy() @ x - b) <= 1e-11 * np.linalg.norm(b)


@pytest.mark.parametrize("A", [O.gen_stencil(5, 20, 20), O.gen_stencil(27, 8, 8, 8),
                               O.gen_stencil(7, 9, 8, 7)])
def test_cg_vs_direct(A):
    xt = O.uniform(7, A.ncols, -1, 1)
    b = O.spmv(A, xt)
    x, k, hist = O.cg(A, b, 10000, 1e-12)
    assert hist[-1] / hist[0] < 1e-24
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-9 * np.linalg.norm(xd)
    # residual history is the recurrence; the true residual must agree
    assert np.linalg.norm(A.to_scipy() @ x - b) <= 1e-10 * np.linalg.norm(b)


def test_cg_quirks():
    A = O.gen_stencil(5, 6, 6)
    # stops exactly at kmax when not converged
    x, k, hist = O.cg(A, O.uniform(1, A.ncols, -1, 1), 3, 1e-14)
    assert k == 3 and len(hist) == 4
    # b = 0: alpha = 0/0 = NaN, never converges, runs to kmax (cg.cpp:62,78)
    x, k, hist = O.cg(A, np.zeros(A.ncols), 5, 1e-8)
    assert k == 5 and np.all(np.isnan(x))


def test_cg_ghost_halo():
    # with a halo, update(p) writes it into p's ghost region every iteration
    A = O.gen_stencil(27, 4, 4, 4, ghost_planes=1)
    b = O.uniform(3, A.nrows, -1, 1)
    x0, k0, _ = O.cg(A, b, 200, 1e-10)  # zero halo: solves the owned block
    Aoo = A.to_scipy()[:, : A.nrows]
    assert np.linalg.norm(Aoo @ x0[: A.nrows] - b) < 1e-9 * np.linalg.norm(b)
    assert np.all(x0[A.nrows:] == 0.0)


def test_cg_threaded_chunked_order():
    """The threaded oracle CG (rows, vector updates and dot products split over
    OpenMP threads; dots summed in a fixed 64-chunk order) gives the same
    result for any thread count, converges in the same number of iterations
    as the sequential oracle with its residual history within 1e-12 and its
    solution within 1e-12 relative, with and without a halo, and keeps the
    closed-form one-iteration convergence on an eigenvector."""
    A = O.gen_stencil(27, 9, 9, 9, ghost_planes=1)
    b = O.uniform(3, A.nrows, -1, 1)
    halo = O.uniform(4, A.ncols - A.nrows, -1, 1)
    for h in (None, halo):
        x1, k1, h1 = O.cg(A, b, 500, 1e-10, halo=h)
        runs = [O.cg(A, b, 500, 1e-10, halo=h, threads=t) for t in (1, 2, 5)]
        for xt, kt, ht in runs:
            assert kt == runs[0][1] and np.array_equal(xt, runs[0][0])
            assert np.array_equal(ht, runs[0][2])
        xt, kt, ht = runs[0]
        assert kt == k1 and k1 > 5
        assert np.max(np.abs(np.sqrt(ht / ht[0]) - np.sqrt(h1 / h1[0]))) <= 1e-12
        assert np.linalg.norm(xt - x1) <= 1e-12 * np.linalg.norm(x1)
    # a Dirichlet sine mode of the 7-point operator is an eigenvector: one
    # iteration, exact solution v / lambda
    n = 6
    A7 = O.gen_stencil(7, n, n, n)
    i = np.arange(n) + 1
    s = np.sin(np.pi * i / (n + 1))
    v = np.einsum("i,j,k->kji", s, s, s).ravel()
    lam = 3 * 2 * (1 - np.cos(np.pi / (n + 1)))
    x, k, _ = O.cg(A7, lam * v, 10, 1e-10, threads=3)
    assert k == 1 and np.allclose(x, v, rtol=0, atol=1e-13)


def test_cg_f32():
    A = O.gen_stencil(5, 16, 16)
    b = O.uniform(2, A.ncols, -1, 1)
    x, k, hist = O.cg_f32(A, A.val.astype(np.float32), b.astype(np.float32), 1000, 1e-5)
    xd = spla.spsolve(A.to_scipy().tocsc(), b)
    assert np.linalg.norm(x - xd) <= 1e-4 * np.linalg.norm(xd)


def test_cg_f32_restarts_meet_true_residual():
    """or_cg_f32_restart (the batched solver's checker): without restarts it
    is the plain FP32 CG (same iterate, count and history); on random SPD
    systems where the plain FP32 CG stops by its recursive residual with a
    true residual ||b - A x|| / ||b|| above rtol (computed here in double
    from the FP32 values), the restarted solve meets rtol, and it restarts
    from x: a few iterations more, not a second solve from zero."""
    from spmv_amd import problems as P
    rtol = 1e-6
    for n, seed in ((3000, 1002), (3000, 1006), (8000, 1000), (8000, 1005)):
        A = P.random_spd(n, seed)
        v32 = A.val.astype(np.float32)
        oc = O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, v32.astype(np.float64))
        b = O.uniform(seed, n, -1, 1).astype(np.float32)

        def true_rel(x):
            r = b.astype(np.float64) - oc.to_scipy() @ x.astype(np.float64)
            return np.linalg.norm(r) / np.linalg.norm(b.astype(np.float64))
        x0, k0, h0 = O.cg_f32(oc, v32, b, 5000, rtol)
        xz, kz, hz = O.cg_f32(oc, v32, b, 5000, rtol, restarts=0)
        assert kz == k0 and np.array_equal(xz, x0) and np.array_equal(hz, h0)
        assert true_rel(x0) > rtol
        x1, k1, _ = O.cg_f32(oc, v32, b, 5000, rtol, restarts=4)
        assert true_rel(x1) <= rtol
        assert k0 < k1 <= k0 + k0 // 2


# ---------------------------------------------------------------- adjoint
def test_adjoint_dense_closed_form():
    n = 4
    ne = O.diffusion_num_edges(n)
    theta = O.uniform(9, ne, 0.5, 2.0)
    b = O.uniform(10, n * n, 0.0, 1.0)
    xt = O.uniform(11, n * n, -1.0, 1.0)
    r = O.adjoint_diffusion(n, theta, b, xt, 1000, 1e-15)
    # dense: A(theta) assembled edge by edge
    def assemble(th):
        A = np.zeros((n * n, n * n))
        nh = n * (n - 1)
        for j in range(n):
            for i in range(n - 1):
                e, a, c = j * (n - 1) + i, i + n * j, i + 1 + n * j
                A[a, a] += th[e]; A[c, c] += th[e]; A[a, c] -= th[e]; A[c, a] -= th[e]
        for j in range(n - 1):
            for i in range(n):
                e, a, c = nh + j * n + i, i + n * j, i + n * (j + 1)
                A[a, a] += th[e]; A[c, c] += th[e]; A[a, c] -= th[e]; A[c, a] -= th[e]
        for t in range(n):
            A[t, t] += th[2 * nh + t]                    # j = 0
            A[t + n * (n - 1), t + n * (n - 1)] += th[2 * nh + n + t]  # j = n-1
            A[n * t, n * t] += th[2 * nh + 2 * n + t]    # i = 0
            A[n - 1 + n * t, n - 1 + n * t] += th[2 * nh + 3 * n + t]  # i = n-1
        return A
    Ad = assemble(theta)
    assert np.allclose(O.gen_diffusion(n, theta).to_scipy().toarray(), Ad, atol=1e-15)
    x = np.linalg.solve(Ad, b)
    lam = np.linalg.solve(Ad.T, x - xt)
    assert np.allclose(r["x"], x, atol=1e-12)
    assert np.allclose(r["lam"], lam, atol=1e-12)
    # dL/dtheta_e = -lam^T (dA/dtheta_e) x
    g = np.empty(ne)
    for e in range(ne):
        d = np.zeros(ne)
        d[e] = 1.0
        g[e] = -lam @ ((assemble(d)) @ x)
    assert np.allclose(r["grad_theta"], g, atol=1e-12)


def test_adjoint_central_differences():
    n = 6
    ne = O.diffusion_num_edges(n)
    theta = np.exp(0.5 * (O.uniform(1, ne, 0, 1) - 0.5))
    b = O.uniform(2, n * n)
    xt = O.uniform(3, n * n, -1, 1)
    r = O.adjoint_diffusion(n, theta, b, xt, 10000, 1e-15)
    eps = 1e-6
    for e in [0, 7, ne // 2, ne - 1]:
        tp, tm = theta.copy(), theta.copy()
        tp[e] += eps
        tm[e] -= eps
        # plain central difference of the loss: limited by cancellation,
        # |error| ~ eps_mach * L / eps + solve error / eps
        Lp = O.adjoint_diffusion(n, tp, b, xt, 10000, 1e-15)["loss"]
        Lm = O.adjoint_diffusion(n, tm, b, xt, 10000, 1e-15)["loss"]
        fd = (Lp - Lm) / (2 * eps)
        assert abs(fd - r["grad_theta"][e]) <= 1e-4 * max(abs(fd), 1e-3)
        # the same central difference evaluated without cancellation:
        # d = x+ - x- solves A(t+) d = -(A(t+) - A(t-)) x-, where in exact
        # arithmetic A(t+) - A(t-) = (t+_e - t-_e) P_e, P_e x = (x_a - x_b)(e_a - e_b)
        # (x_a e_a for a boundary edge); L+ - L- = d.(x- - xt) + d.d/2 exactly
        Ap = O.gen_diffusion(n, tp).to_scipy().toarray()
        Am = O.gen_diffusion(n, tm).to_scipy().toarray()
        xm = np.linalg.solve(Am, b)
        a, c = diffusion_edge_nodes(n, e)
        Pe_x = np.zeros(n * n)
        if c is None:
            Pe_x[a] = xm[a]
        else:
            Pe_x[a], Pe_x[c] = xm[a] - xm[c], xm[c] - xm[a]
        dt = tp[e] - tm[e]
        d = np.linalg.solve(Ap, -dt * Pe_x)
        fd_exact = (d @ (xm - xt) + 0.5 * d @ d) / dt
        assert abs(fd_exact - r["grad_theta"][e]) <= 1e-6 * abs(fd_exact)


def diffusion_edge_nodes(n, e):
    """(a, b) node pair of edge e in the numbering of or_gen_diffusion; b is
    None for a boundary edge."""
    nh = n * (n - 1)
    if e < nh:
        j, i = divmod(e, n - 1)
        return i + n * j, i + 1 + n * j
    if e < 2 * nh:
        j, i = divmod(e - nh, n)
        return i + n * j, i + n * (j + 1)
    side, t = divmod(e - 2 * nh, n)
    return [t, t + n * (n - 1), n * t, n - 1 + n * t][side], None


# ---------------------------------------------------------------- complex values
@pytest.mark.parametrize("A", [O.gen_stencil(7, 9, 8, 7), O.gen_stencil(27, 5, 5, 5, ghost_planes=1),
                               O.gen_demo1d(50)], ids=["poisson7", "stencil27_ghost", "demo1d"])
def test_complex_spmv_bitexact_vs_scipy(A):
    """Matrix<std::complex<T>> (Matrix.cpp:441-444): the row order of the
    real product with std::complex's product; scipy's csr_matvec evaluates
    the same formula in the same order, so the results are identical."""
    import scipy.sparse as sp
    v = A.val + 1j * O.uniform(3, A.nnz, -1, 1)
    x = O.uniform(4, A.ncols, -1, 1) + 1j * O.uniform(5, A.ncols, -1, 1)
    S = sp.csr_matrix((v, A.colidx, A.rowptr), shape=(A.nrows, A.ncols))
    y = O.spmv_complex(A, v, x)
    assert np.array_equal(y, S @ x)
    assert np.allclose(y, S.toarray() @ x, rtol=1e-13, atol=1e-13)
    v32, x32 = v.astype(np.complex64), x.astype(np.complex64)
    S32 = sp.csr_matrix((v32, A.colidx, A.rowptr), shape=(A.nrows, A.ncols))
    assert np.array_equal(O.spmv_complex(A, v32, x32), S32 @ x32)
    xt = x[:A.nrows]
    assert np.array_equal(O.spmv_t_complex(A, v, xt), S.T @ xt)
    # a real matrix as complex: the real part is the real product
    assert np.array_equal(O.spmv_complex(A, A.val.astype(np.complex128), x.real + 0j).real,
                          O.spmv(A, x.real))
