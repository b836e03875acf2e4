"""Known-answer vectors (tests/golden/spmv/*.json, written by
tests/golden/make_spmv_golden.py and checked there against scipy) replayed
on the CPU: the generators rebuild the same matrices, the oracle and scipy
reproduce y bit-for-bit, and the oracle reproduces the CG iteration count and
residual history exactly."""
import glob
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = sorted(glob.glob(os.path.join(HERE, "golden", "spmv", "*.json")))
GEN = {
    "demo1d_64": lambda: O.gen_demo1d(64),
    "poisson5_8x8": lambda: O.gen_stencil(5, 8, 8),
    "poisson7_6x6x6": lambda: O.gen_stencil(7, 6, 6, 6),
    "stencil27_5x5x5": lambda: O.gen_stencil(27, 5, 5, 5),
    "stencil27_5x5x5_ghost": lambda: O.gen_stencil(27, 5, 5, 5, ghost_planes=1),
}


def unhex(a):
    return np.array([float.fromhex(v) for v in a])


def load(path):
    with open(path) as f:
        return json.load(f)


def test_fixture_set_complete():
    assert sorted(os.path.basename(f)[:-5] for f in FILES) == sorted(GEN)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-5] for f in FILES])
def test_golden_spmv_and_cg(path):
    g = load(path)
    A = GEN[g["case"]]()
    assert (A.nrows, A.ncols, int(A.nnz)) == (g["nrows"], g["ncols"], g["nnz"])
    assert int(np.sum(A.rowptr, dtype=np.int64)) == g["rowptr_sum"]
    assert int(np.sum(A.colidx, dtype=np.int64)) == g["colidx_sum"]
    assert float(np.sum(A.val)).hex() == g["val_sum"]
    x = O.uniform(g["x_seed"], A.ncols, *g["x_range"])
    y = unhex(g["y"])
    assert np.array_equal(O.spmv(A, x), y)
    S = sp.csr_matrix((A.val, A.colidx, A.rowptr), shape=(A.nrows, A.ncols))
    assert np.array_equal(S @ x, y)
    if "cg_iterations" in g:
        b = unhex(g["b"])
        assert np.array_equal(O.spmv(A, O.uniform(g["xt_seed"], A.ncols, -1.0, 1.0)), b)
        xs, k, hist = O.cg(A, b, 10000, g["cg_rtol"])
        assert k == g["cg_iterations"]
        assert np.array_equal(hist[:k + 1], unhex(g["cg_history"]))
        assert np.array_equal(xs, unhex(g["cg_x"]))
