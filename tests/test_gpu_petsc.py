"""PETSc binary input on the GPU: the matrix the reader builds (C ABI
spmvx_petsc_read_matrix, the Python wrapper and the C++ cgmain demo) gives
the oracle's products bit-for-bit and the oracle's CG."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import torch

import oracle as O
import oracle.petsc as OP
import spmv_amd as S
from spmv_amd import problems as P

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spmv_amd", "bin")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to(DEV)


def _write_poisson(path, n):
    A = P.stencil(5, n, n)
    OP.write_matrix(path, A.nrows, A.ncols, A.rowptr, A.colidx, A.val)
    return A


def test_c_abi_read_matrix_whole(tmp_path):
    path = str(tmp_path / "A.dat")
    A = _write_poisson(path, 48)
    h, cm, rm = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    st = S.lib().spmvx_petsc_read_matrix(path.encode(), 1, 0, S.F64, ctypes.byref(h),
                                         ctypes.byref(cm), ctypes.byref(rm))
    assert st == 0, S.lib().spmvx_last_error()
    M = S.Matrix(None, None, None, _handle=h)  # owns h from here on
    try:
        x = P.uniform(4, A.ncols, -1, 1)
        oc = O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val)
        assert np.array_equal(M.mult(_dev(x)).cpu().numpy(), O.spmv(oc, x))
        # same matrix, same kernels: CG identical to the generated matrix's
        b = _dev(P.uniform(5, A.nrows, -1, 1))
        x1, k1 = S.cg(M, b, 5000, 1e-10)
        G = S.Matrix(A.rowptr, A.colidx, A.val, ncols=A.ncols)
        x2, k2 = S.cg(G, b, 5000, 1e-10)
        assert k1 == k2 and torch.equal(x1, x2)
    finally:
        del M
        S.lib().spmvx_l2gmap_destroy(cm)
        S.lib().spmvx_l2gmap_destroy(rm)


@pytest.mark.parametrize("nranks,rank", [(3, 0), (3, 1), (4, 3)])
def test_read_matrix_rank_block(tmp_path, nranks, rank):
    path = str(tmp_path / "A.dat")
    _write_poisson(path, 40)
    o = OP.read_matrix(path, nranks, rank)
    M = S.read_petsc_matrix(path, nranks, rank)
    ncl, ng = o["ncols_local"], len(o["ghosts"])
    assert (M.nrows, M.ncols) == (o["nrows_local"], ncl + ng)
    cm = M.col_map()
    assert cm.local_size() == ncl and cm.num_ghosts() == ng
    assert list(cm.ghosts()) == list(o["ghosts"])
    # ghost update from a host halo, then the product of the local block
    xl = P.uniform(6, ncl + ng, -1, 1)
    halo = P.uniform(7, ng, -1, 1)
    cm.set_halo(halo)
    xd = _dev(xl)
    cm.update(xd)
    xl[ncl:] = halo
    assert np.array_equal(xd.cpu().numpy(), xl)
    oc = O.CSR(o["nrows_local"], ncl + ng, o["rowptr"].astype(np.int32),
               o["colidx"].astype(np.int32), o["values"])
    assert np.array_equal(M.mult(xd).cpu().numpy(), O.spmv(oc, xl))


def test_cgmain_petsc_files(tmp_path):
    exe = os.path.join(BIN, "cgmain")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} was not built")
    A = _write_poisson(str(tmp_path / "petsc_mat7.dat"), 64)
    b = P.uniform(8, A.nrows, -1, 1)
    OP.write_vector(str(tmp_path / "petsc_vec7.dat"), b)
    out = subprocess.run([exe, "petsc", "7"], capture_output=True, text=True, timeout=120,
                         cwd=str(tmp_path))
    assert out.returncode == 0, out.stderr
    text = out.stdout
    assert "Read file: petsc_mat7.dat: 4096x4096 = %d" % A.nnz in text
    assert "Global vec size = 4096" in text
    m = re.search(r"^x.norm = (\S+) in (\d+) iterations$", text, re.M)
    x, k, _ = O.cg(O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val), b, 10000, 1e-10)
    assert abs(int(m.group(2)) - k) <= 1
    assert abs(float(m.group(1)) - np.linalg.norm(x)) <= 1e-9 * np.linalg.norm(x)
    rn = float(re.search(r"^r.norm = (\S+)$", text, re.M).group(1))
    assert rn <= 2e-10 * np.linalg.norm(b)
