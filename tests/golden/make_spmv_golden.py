"""Writes tests/golden/spmv/*.json: small SpMV and CG known-answer vectors of
the reference's demo and stencil operators, computed by the C oracle (the
sequential per-row order of spmv/Matrix.cpp:199-204 and the loop of
spmv/cg.cpp:11-83) and checked here against scipy before they are written:
y bit-for-bit against scipy's csr_matvec (same sequential row order), the CG
solution against scipy's direct solve.

Doubles are stored as hex strings (float.hex) so they round-trip exactly.
x and x_true come from the splitmix64 streams of oracle.uniform.

    python tests/golden/make_spmv_golden.py
"""
import json
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402

# name -> (generator call, CG rtol or None)
CASES = {
    "demo1d_64": (lambda: O.gen_demo1d(64), 1e-10),
    "poisson5_8x8": (lambda: O.gen_stencil(5, 8, 8), 1e-10),
    "poisson7_6x6x6": (lambda: O.gen_stencil(7, 6, 6, 6), 1e-10),
    "stencil27_5x5x5": (lambda: O.gen_stencil(27, 5, 5, 5), 1e-10),
    "stencil27_5x5x5_ghost": (lambda: O.gen_stencil(27, 5, 5, 5, ghost_planes=1), None),
}
X_SEED, XT_SEED = 101, 202


def hexs(a):
    return [float(v).hex() for v in np.asarray(a, dtype=np.float64)]


def scipy_of(A):
    return sp.csr_matrix((A.val, A.colidx, A.rowptr), shape=(A.nrows, A.ncols))


def main():
    out = os.path.join(HERE, "spmv")
    os.makedirs(out, exist_ok=True)
    for name, (gen, rtol) in CASES.items():
        A = gen()
        x = O.uniform(X_SEED, A.ncols, -1.0, 1.0)
        y = O.spmv(A, x)
        ys = scipy_of(A) @ x
        assert np.array_equal(y, ys), name
        rec = {"case": name, "nrows": A.nrows, "ncols": A.ncols, "nnz": int(A.nnz),
               "rowptr_sum": int(np.sum(A.rowptr, dtype=np.int64)),
               "colidx_sum": int(np.sum(A.colidx, dtype=np.int64)),
               "val_sum": float(np.sum(A.val)).hex(),
               "x_seed": X_SEED, "x_range": [-1.0, 1.0], "y": hexs(y)}
        if rtol is not None:
            xt = O.uniform(XT_SEED, A.ncols, -1.0, 1.0)
            b = O.spmv(A, xt)
            xs, k, hist = O.cg(A, b, 10000, rtol)
            direct = spla.spsolve(scipy_of(A).tocsc(), b)
            rel = np.linalg.norm(xs[:A.nrows] - direct) / np.linalg.norm(direct)
            assert rel < 1e-7, (name, rel)
            rec.update({"xt_seed": XT_SEED, "b": hexs(b), "cg_rtol": rtol, "cg_iterations": int(k),
                        "cg_history": hexs(hist[:k + 1]), "cg_x": hexs(xs)})
        with open(os.path.join(out, name + ".json"), "w") as f:
            json.dump(rec, f, indent=0)
        print(name, A.nrows, A.nnz, rec.get("cg_iterations"), file=sys.stderr)


if __name__ == "__main__":
    main()
