"""The C++ drop-in API (spmv::Matrix, spmv::L2GMap, spmv::cg) through the
demo drivers built from spmv_amd/demos/, checked against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from spmv_amd import problems as P

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "spmv_amd", "bin")


def run(exe, *args):
    path = os.path.join(BIN, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} was not built")
    out = subprocess.run([path, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout


def field(text, name):
    return float(re.search(rf"^{re.escape(name)} = (\S+)$", text, re.M).group(1))


def test_matvecmain_matches_oracle():
    N, n_apply = 20000, 50
    out = run("matvecmain", N, n_apply)
    A = O.gen_demo1d(N)
    z = np.arange(N, dtype=np.float64) / N
    v = np.exp(-10 * np.power(5 * (z - 0.5), 2.0))
    for _ in range(n_apply):
        v = O.spmv(A, v)
    want = float(np.sum(v * v))
    # the SpMVs are bit-identical; the initial exp/pow may differ by an ulp
    assert abs(field(out, "norm") - want) <= 1e-12 * want


def test_cgmain_matches_oracle():
    n = 128
    out = run("cgmain", n, 1e-8, 10000, 1)
    A = O.gen_stencil(5, n, n)
    xt = O.uniform(1, n * n, -1, 1)
    b = O.spmv(A, xt)
    x, k, _ = O.cg(A, b, 10000, 1e-8)
    assert abs(field(out, "iterations") - k) <= 1
    assert field(out, "rel.residual") <= 1e-7
    assert abs(field(out, "x.norm") - np.linalg.norm(x)) <= 1e-8 * np.linalg.norm(x)


@pytest.mark.parametrize("mode", ["real", "complex"])
def test_restrictmain_petsc(tmp_path, mode):
    """demos/restrictmain.cpp's loop (transpmult, reverse_update, update,
    product) through the C++ API on a PETSc file, real and
    Matrix<std::complex<double>>, against dense numpy."""
    import oracle.petsc as OP
    rng = np.random.default_rng(9)
    nr, nc = 40, 90
    rows = [np.sort(rng.choice(nc, 6, replace=False)) for _ in range(nr)]
    rowptr = np.concatenate([[0], np.cumsum([len(r) for r in rows])])
    colidx = np.concatenate(rows)
    val = P.uniform(31, len(colidx), -0.3, 0.3)
    OP.write_matrix(str(tmp_path / "R.dat"), nr, nc, rowptr, colidx, val)
    b = P.uniform(32, nr, -1, 1)
    OP.write_vector(str(tmp_path / "b.dat"), b)
    args = [str(tmp_path / "R.dat"), str(tmp_path / "b.dat")] + (["complex"] if mode == "complex" else [])
    out = run("restrictmain", *args)
    R = np.zeros((nr, nc), complex)
    for i in range(nr):
        R[i, rows[i]] = val[rowptr[i]:rowptr[i + 1]]
    q = b.astype(complex)
    if mode == "complex":
        R, q = R * (1 + 0.5j), q * (1 - 0.25j)
    for _ in range(10):
        p = R.T @ q
        pn = float(np.sum(np.abs(p) ** 2))
        q = R @ p
        qn = float(np.sum(np.abs(q) ** 2))
    assert field(out, "mat.equal") == 1
    assert abs(field(out, "pnorm") - pn) <= 1e-12 * pn
    assert abs(field(out, "qnorm") - qn) <= 1e-12 * qn


def test_create_matrix_one_rank_of_two():
    """Matrix::create_matrix (Matrix.cpp:213-438) for rank 0 of a two-rank
    decomposition: ghost columns renumbered in ascending global order (a
    repeated ghost merged), entries of each row sorted by column with
    repeated columns summed in stored order; the column map carries the new
    ghosts; a call with ghost rows is refused (one process drives one rank)."""
    out = run("createmain")
    rows_in = [[(1, 2.0), (0, 4.0), (6, -1.0), (4, -0.5)], [(1, 3.0), (1, 0.25)],
               [(5, 7.0), (2, -2.0), (3, 1.5)], [(3, 6.0), (2, -3.0)]]
    col_ghosts = [7, 5, 7]
    ghosts = sorted(set(col_ghosts))
    glob = lambda c: c if c < 4 else col_ghosts[c - 4]
    new = lambda g: g if g < 4 else 4 + ghosts.index(g)
    outer, inner, values = [0], [], []
    for row in rows_in:
        acc = {}
        for c, v in row:  # stored order: duplicates summed in that order
            k = new(glob(c))
            acc[k] = acc[k] + v if k in acc else v
        for k in sorted(acc):
            inner.append(k)
            values.append(acc[k])
        outer.append(len(inner))

    def ints(name):
        return [int(t) for t in re.search(rf"^{name} =(.*)$", out, re.M).group(1).split()]

    def floats(name):
        return [float(t) for t in re.search(rf"^{name} =(.*)$", out, re.M).group(1).split()]
    assert field(out, "rows") == 4 and field(out, "cols") == 4 + len(ghosts)
    assert ints("outer") == outer and ints("inner") == inner and floats("values") == values
    assert ints("ghosts") == ghosts
    y = [sum(values[k] * (inner[k] + 1) for k in range(outer[r], outer[r + 1])) for r in range(4)]
    assert floats("y") == y
    assert "row_ghosts must be empty" in re.search(r"^error = (.*)$", out, re.M).group(1)
