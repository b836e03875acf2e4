"""spmvx_cg_noalloc (CG with every buffer in the library's fixed device
scratch) against spmvx_cg, through the C ABI: the same kernels in the same
order, so x, the iteration count and the residual history must be
bit-identical.  Two solves in a row on one GPU share the scratch; the second
is smaller, so stale contents of the first would show."""
import ctypes

import numpy as np
import pytest
import torch

import oracle as O
import spmv_amd as S
from spmv_amd import problems as P

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(DEV)


def _solve_both(M, b, kmax, rtol):
    x_ref, k_ref, h_ref = S.cg(M, b, kmax, rtol, history=True)
    x = torch.full((M.ncols,), float("nan"), dtype=torch.float64, device=DEV)
    x_na, k_na, h_na = S.cg_noalloc(M, b, kmax, rtol, x=x, history=True)
    assert x_na.data_ptr() == x.data_ptr()
    assert k_na == k_ref and k_ref > 10
    assert torch.equal(x_na, x_ref)
    assert np.array_equal(h_na, h_ref)
    return x_na, k_na, h_na


def test_noalloc_cg_matches_cg_two_calls():
    # first: a 27-point block with its ghost plane and a halo of zeros,
    # through the column map (ghost columns, pack indices)
    A = P.stencil(27, 24, 24, 24, ghost_planes=1)
    m = S.L2GMap(**P.slab_decomposition(24, 24, 24))
    M = S.Matrix(A.rowptr, A.colidx, A.val, col_map=m)
    b = P.uniform(5, A.nrows, -1, 1)
    x, k, hist = _solve_both(M, _dev(b), 2000, 1e-10)
    xo, ko, histo = O.cg(O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val), b, 2000, 1e-10)
    assert abs(k - ko) <= 1
    mm = min(k, ko) + 1
    assert np.max(np.abs(np.sqrt(hist[:mm] / hist[0]) - np.sqrt(histo[:mm] / histo[0]))) <= 1e-10
    # second, smaller: a 5-point Poisson system without ghost columns
    A2 = P.stencil(5, 40, 40)
    M2 = S.Matrix(A2.rowptr, A2.colidx, A2.val, ncols=A2.ncols)
    b2 = P.uniform(6, A2.nrows, -1, 1)
    _solve_both(M2, _dev(b2), 2000, 1e-10)
    # and the sliced layout of the first matrix
    M.set_kernel(S.SLICED_VARIANT)
    _solve_both(M, _dev(b), 2000, 1e-10)


def test_noalloc_cg_rejects_what_does_not_fit():
    """A solve needing more than the scratch, or a column map whose device
    plan has not been built, is refused before any kernel runs."""
    n = 1_500_000  # 8 * 3n bytes > 32 MiB
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    M = S.Matrix(rp, ci, np.full(n, 2.0), ncols=n)
    b = torch.ones(n, dtype=torch.float64, device=DEV)
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    with pytest.raises(S.SpmvError, match="scratch"):
        S.cg_noalloc(M, b, 10, 1e-10, x=x)
    assert torch.all(x == 0)
    # the raw C entry point reports SPMVX_ERR_ARG (1)
    its = ctypes.c_int32(-1)
    st = S.lib().spmvx_cg_noalloc(M.handle, None, ctypes.c_void_p(b.data_ptr()),
                                  ctypes.c_void_p(x.data_ptr()), 10, 1e-10, ctypes.byref(its),
                                  None, None)
    assert st == 1 and its.value == -1
    A = P.stencil(27, 8, 8, 8, ghost_planes=1)
    m = S.L2GMap(**P.slab_decomposition(8, 8, 8))
    M2 = S.Matrix(A.rowptr, A.colidx, A.val, col_map=m)
    b2 = _dev(P.uniform(7, A.nrows, -1, 1))
    with pytest.raises(S.SpmvError, match="device plan"):
        S.cg_noalloc(M2, b2, 100, 1e-10)
    # after one update the plan exists and the solve runs
    v = torch.zeros(A.ncols, dtype=torch.float64, device=DEV)
    m.update(v)
    S.cg_noalloc(M2, b2, 100, 1e-10)
