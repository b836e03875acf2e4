"""PETSc binary input (spmv/read_petsc.cpp): the oracle's writer pinned by
hand-assembled bytes, the oracle's reader by round trips and by the
reassembly of all ranks' blocks, and the library's host parser
(spmvx_petsc_csr_*) checked against the oracle for 1..4-way splits,
including the exchange plan its ghosts produce.  CPU only."""
import os
import struct

import numpy as np
import pytest

import oracle.l2gmap as OL
import oracle.petsc as OP
import spmv_amd as S
from spmv_amd import problems as P


def _poisson(n):
    A = P.stencil(5, n, n)
    return A.nrows, A.ncols, A.rowptr.astype(np.int64), A.colidx.astype(np.int64), A.val


def _random_rect(nrows, ncols, seed, shuffle=True):
    rng = np.random.default_rng(seed)
    rp, ci = [0], []
    for i in range(nrows):
        k = 0 if i % 9 == 4 else int(rng.integers(1, 12))
        cols = rng.choice(ncols, size=min(k, ncols), replace=False)
        if not shuffle:
            cols = np.sort(cols)
        ci.extend(cols.tolist())
        rp.append(len(ci))
    vals = P.uniform(seed, len(ci), -1, 1)
    return nrows, ncols, np.array(rp, np.int64), np.array(ci, np.int64), vals


CASES = {
    "poisson5_12": lambda: _poisson(12),
    "rect_unsorted": lambda: _random_rect(37, 53, 3),
    "rect_wide": lambda: _random_rect(20, 90, 4, shuffle=False),
    "square_unsorted": lambda: _random_rect(64, 64, 5),
}


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("petsc")
    out = {}
    for name, gen in CASES.items():
        nr, nc, rp, ci, v = gen()
        path = str(d / f"{name}.dat")
        OP.write_matrix(path, nr, nc, rp, ci, v)
        out[name] = (path, (nr, nc, rp, ci, v))
    return out


def test_writer_known_bytes(tmp_path):
    # [[1.5, 0, -2], [0, 4, 0]] with the first row's columns written as (2, 0)
    path = str(tmp_path / "m.dat")
    OP.write_matrix(path, 2, 3, [0, 2, 3], [2, 0, 1], [-2.0, 1.5, 4.0])
    want = (struct.pack(">4i", 1211216, 2, 3, 3) + struct.pack(">2i", 2, 1)
            + struct.pack(">3i", 2, 0, 1) + struct.pack(">3d", -2.0, 1.5, 4.0))
    assert open(path, "rb").read() == want
    vpath = str(tmp_path / "v.dat")
    OP.write_vector(vpath, [0.25, -1.0, 3.0])
    assert open(vpath, "rb").read() == struct.pack(">2i", 1211214, 3) + struct.pack(">3d", 0.25, -1.0, 3.0)
    # both readers sort the row by column
    for r in (OP.read_matrix(path), S.read_petsc_csr(path)):
        assert list(r["rowptr"]) == [0, 2, 3]
        assert list(r["colidx"]) == [0, 2, 1]
        assert list(r["values"]) == [1.5, -2.0, 4.0]
        assert len(r["ghosts"]) == 0


def test_owner_ranges():
    assert OP.owner_ranges(4, 10) == [0, 3, 6, 8, 10]
    assert OP.owner_ranges(3, 2) == [0, 1, 2, 2]
    assert list(S._owner_ranges(4, 10)) == [0, 3, 6, 8, 10]


def _dense_global(blocks, nrows, ncols):
    """Reassemble the global matrix from every rank's block."""
    D = np.zeros((nrows, ncols))
    for b in blocks:
        r0 = b["row_ranges"][b["rank"]]
        c0 = b["col_ranges"][b["rank"]]
        ncl = b["ncols_local"]
        for i in range(b["nrows_local"]):
            for k in range(b["rowptr"][i], b["rowptr"][i + 1]):
                lc = b["colidx"][k]
                g = c0 + lc if lc < ncl else b["ghosts"][lc - ncl]
                D[r0 + i, g] += b["values"][k]
    return D


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("nranks", [1, 2, 3, 4])
def test_library_parser_matches_oracle(files, name, nranks):
    path, (nr, nc, rp, ci, v) = files[name]
    blocks = []
    ghosts = []
    for rank in range(nranks):
        o = OP.read_matrix(path, nranks, rank)
        c = S.read_petsc_csr(path, nranks, rank)
        assert (c["nrows"], c["ncols"]) == (nr, nc)
        assert (c["nrows_local"], c["ncols_local"]) == (o["nrows_local"], o["ncols_local"])
        for k in ("rowptr", "colidx", "ghosts"):
            assert np.array_equal(c[k], o[k]), k
        assert np.array_equal(c["values"], o["values"])
        for q in range(nranks):
            assert np.array_equal(c["remote"][q], o["remote"][q])
        # rows ascending in local column: owned columns first, ghosts after
        for i in range(o["nrows_local"]):
            seg = c["colidx"][c["rowptr"][i]:c["rowptr"][i + 1]]
            assert np.all(np.diff(seg) > 0)
        o["rank"] = rank
        blocks.append(o)
        ghosts.append(list(o["ghosts"]))
    # reassembling every rank's block gives the file's matrix back
    D = np.zeros((nr, nc))
    for i in range(nr):
        for k in range(rp[i], rp[i + 1]):
            D[i, ci[k]] += v[k]
    assert np.array_equal(_dense_global(blocks, nr, nc), D)
    # the remote lists are the other ranks' ghosts in this rank's range
    cr = OP.owner_ranges(nranks, nc)
    for rank in range(nranks):
        for q in range(nranks):
            want = [g for g in ghosts[q] if cr[rank] <= g < cr[rank + 1]] if q != rank else []
            assert list(blocks[rank]["remote"][q]) == want


@pytest.mark.parametrize("nranks", [2, 3, 4])
def test_exchange_plan_of_read_matrix(files, nranks):
    """The column map built from the parse (C ABI, host only) has the plan the
    reference's MPI constructor computes (oracle/l2gmap.py restatement)."""
    path, (nr, nc, rp, ci, v) = files["poisson5_12"]
    cr = OP.owner_ranges(nranks, nc)
    sizes = [cr[q + 1] - cr[q] for q in range(nranks)]
    parsed = [S.read_petsc_csr(path, nranks, r) for r in range(nranks)]
    plans = OL.build_plans(sizes, [list(p["ghosts"]) for p in parsed])
    for r, p in enumerate(parsed):
        m = S.L2GMap(p["ncols_local"], p["ghosts"], local_sizes=sizes, rank=r,
                     remote_ghosts=p["remote"])
        got = m.plan()
        for k in ("neighbours", "send_count", "recv_count", "send_offset", "recv_offset",
                  "indexbuf"):
            assert got[k] == plans[r][k], (k, r)


def test_vector_split(tmp_path):
    path = str(tmp_path / "v.dat")
    vals = P.uniform(9, 23, -1, 1)
    OP.write_vector(path, vals)
    for nranks in (1, 2, 5):
        parts = [S.read_petsc_vector(path, nranks, r) for r in range(nranks)]
        for r in range(nranks):
            assert np.array_equal(parts[r], OP.read_vector(path, nranks, r))
        assert np.array_equal(np.concatenate(parts), vals)


def test_errors(tmp_path, files):
    with pytest.raises(S.SpmvError, match="Could not open file"):
        S.read_petsc_csr(str(tmp_path / "missing.dat"))
    with pytest.raises(S.SpmvError, match="Could not open file"):
        S.read_petsc_vector(str(tmp_path / "missing.dat"))
    vpath = str(tmp_path / "v.dat")
    OP.write_vector(vpath, [1.0, 2.0])
    with pytest.raises(S.SpmvError, match="Bad signature in PETSc Matrix file"):
        S.read_petsc_csr(vpath)
    mpath = files["poisson5_12"][0]
    with pytest.raises(S.SpmvError, match="Bad signature in PETSc Vector file"):
        S.read_petsc_vector(mpath)
    data = open(mpath, "rb").read()
    cut = str(tmp_path / "cut.dat")
    open(cut, "wb").write(data[:-8])
    with pytest.raises(S.SpmvError, match="truncated"):
        S.read_petsc_csr(cut)
    bad = str(tmp_path / "badcol.dat")
    OP.write_matrix(bad, 2, 2, [0, 1, 2], [0, 5], [1.0, 1.0])
    with pytest.raises(S.SpmvError, match="column index out of range"):
        S.read_petsc_csr(bad)
    dup = str(tmp_path / "dup.dat")
    OP.write_matrix(dup, 2, 2, [0, 2, 3], [1, 1, 0], [1.0, 2.0, 3.0])
    with pytest.raises(S.SpmvError, match="duplicate column"):
        S.read_petsc_csr(dup)
    with pytest.raises(S.SpmvError, match="rank"):
        S.read_petsc_csr(mpath, 2, 2)
