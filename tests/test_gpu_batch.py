"""GPU checks of the batched persistent CG (extension 2) against the oracle's
FP32 CG (the loop of spmv/cg.cpp:50-80 in single precision, with the batched
solver's true-residual check: when the recursive residual meets rtol, b - A x
is formed in double and CG restarts from x if it does not), per system and
right-hand side.  Every right-hand side must reach a true residual
||b - A x|| / ||b|| <= rtol = 1e-6, computed on the host in double, unless
even the exact solution rounded to FP32 does not (then within 1.5 times that
floor)."""
import numpy as np
import pytest
import torch

import oracle as O
import spmv_amd as S
from spmv_amd import autotune as T
from spmv_amd import problems as P

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def csr32(A):
    """The system the batched solver solves: the values rounded to FP32 (the
    true residuals below are taken against it, in double)."""
    return O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx,
                 A.val.astype(np.float32).astype(np.float64))


def fp32_floor(oc, b32):
    """||b - A fl32(x*)|| / ||b||: the true residual of the exact solution
    rounded to FP32 (computed by the oracle's FP64 CG to 1e-15), below which
    no solver returning FP32 x can go; for these systems it is ~1e-7 to 1e-6."""
    b = b32.astype(np.float64)
    xs, _, _ = O.cg(oc, b, 20000, 1e-15, threads=8)
    r = b - O.spmv(oc, xs[: oc.nrows].astype(np.float32).astype(np.float64))
    return float(np.linalg.norm(r) / np.linalg.norm(b))


def meets_rtol(rel, oc, b32, rtol=1e-6):
    """The batched solver's accuracy contract: the true residual meets rtol,
    or -- when even the FP32-rounded exact solution does not meet rtol with
    some room -- stays within 1.5 times that FP32 floor."""
    return rel <= rtol or rel <= 1.5 * fp32_floor(oc, b32)


def make_systems(sizes, seed):
    return [P.random_spd(int(n), seed + i) for i, n in enumerate(sizes)]


def rhs(nrows, nrhs, seed):
    return P.uniform(seed, nrows * nrhs, -1, 1).astype(np.float32).reshape(nrows, nrhs)


@pytest.mark.parametrize("reorder", [False, True], ids=["caller_order", "rcm"])
@pytest.mark.parametrize("nrhs,rpm", [(4, 0), (4, 700), (2, 1000), (1, 16384)])
def test_batch_matches_oracle(nrhs, rpm, reorder, monkeypatch):
    """Without reordering the solver sums in another order than the oracle;
    with the reverse Cuthill-McKee order (opt-in) the system is the same
    matrix with permuted rows and columns, whose FP32 CG rounds differently
    again.  Either way the iteration counts agree within 8 or 12% (one
    true-residual restart more or less), the true residual meets rtol and the
    solutions agree."""
    monkeypatch.setenv("SPMVX_BATCH_REORDER", "1" if reorder else "0")
    systems = make_systems([1500, 700, 3000, 2200, 999, 4100, 1800, 640], 100)
    B = S.Batch(systems, nrhs=nrhs)
    B.set_config(1024, 0, rpm)
    b = rhs(B.nrows, nrhs, 7)
    x, its = B.solve(torch.from_numpy(b).to(DEV), max_its=2000, rtol=1e-6)
    xh = x.cpu().numpy()
    for s, A in enumerate(systems):
        lo, hi = B.sys_row[s], B.sys_row[s + 1]
        oc = csr32(A)
        for j in range(nrhs):
            xo, ko, _ = O.cg_f32(oc, A.val.astype(np.float32), b[lo:hi, j], 2000, 1e-6,
                                 restarts=4)
            # the reduction orders differ (+-3 without restarts), and so may
            # whether a true-residual restart is needed: a restart begins a new
            # Krylov space and asks for a four times lower recursive residual,
            # up to about a tenth of the iterations more
            tol = max(8, int(0.12 * ko))
            assert abs(int(its[s, j]) - ko) <= tol, (s, j, its[s, j], ko)
            bn = np.linalg.norm(b[lo:hi, j])
            r = b[lo:hi, j].astype(np.float64) - O.spmv(oc, xh[lo:hi, j].astype(np.float64))
            assert meets_rtol(np.linalg.norm(r) / bn, oc, b[lo:hi, j]), (s, j)
            assert np.linalg.norm(xh[lo:hi, j] - xo) <= 1e-4 * np.linalg.norm(xo)


def check_true_residuals(systems, B, b, x):
    """Every system and right-hand side meets the accuracy contract."""
    for s, A in enumerate(systems):
        lo, hi = B.sys_row[s], B.sys_row[s + 1]
        oc = csr32(A)
        for j in range(b.shape[1]):
            r = b[lo:hi, j].astype(np.float64) - O.spmv(oc, x[lo:hi, j].astype(np.float64))
            assert meets_rtol(np.linalg.norm(r) / np.linalg.norm(b[lo:hi, j]), oc, b[lo:hi, j]), \
                (s, j)


def test_batch_every_config_and_replay():
    """Every (workgroup size, lanes per row) pair converges in about the same
    number of iterations (within 8 or 12%: the reduction order differs, and
    with it whether a true-residual restart is needed) to a solution whose
    true residual meets the tolerance."""
    systems = make_systems([900, 2600, 1300, 300, 4000], 300)
    B = S.Batch(systems, nrhs=4)
    bh = rhs(B.nrows, 4, 3)
    bd = torch.from_numpy(bh).to(DEV)
    ref, its0 = B.solve(bd, max_its=1000, rtol=1e-6)
    for bs in T.BATCH_BLOCKS:
        for g in T.BATCH_LANES + (1, 16):
            for rpm in (0, 300, 1000):  # 0: one workgroup per system; else teams
                B.set_config(bs, g, rpm)
                x, its = B.solve(bd, max_its=1000, rtol=1e-6)
                assert np.max(np.abs(its - its0)) <= max(8, int(0.12 * its0.max()))
                check_true_residuals(systems, B, bh, x.cpu().numpy())
    # graph replays: same answer every time, with eager work in between
    B.set_config(256, 0, 500)
    out = torch.empty_like(bd)
    first, _ = B.solve(bd, out, max_its=1000, rtol=1e-6)
    first = first.clone()
    for _ in range(3):
        torch.cuda.synchronize()
        _ = (bd * 2.0).sum().item()
        x, _ = B.solve(bd, out, max_its=1000, rtol=1e-6)
        assert torch.equal(x, first)


def test_batch_max_its_and_zero_rhs():
    systems = make_systems([800, 1200], 50)
    B = S.Batch(systems, nrhs=4)
    b = rhs(B.nrows, 4, 1)
    b[: systems[0].nrows, 1] = 0.0
    x, its = B.solve(torch.from_numpy(b).to(DEV), max_its=3, rtol=1e-12)
    assert np.all(its[:, [0, 2, 3]] == 3)
    x0, its0 = B.solve(torch.from_numpy(b).to(DEV), max_its=0, rtol=1e-6)
    assert np.all(its0 == 0) and torch.all(x0 == 0)


def test_autotuner_caches(tmp_path, monkeypatch):
    monkeypatch.setenv("SPMV_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    systems = make_systems([2000, 1000, 500], 9)
    B = S.Batch(systems, nrhs=4)
    bd = torch.from_numpy(rhs(B.nrows, 4, 2)).to(DEV)
    best, table = T.tune_batch(B, bd, max_its=500, rtol=1e-6, reps=1)
    assert table is not None
    assert len(table) == len(T.BATCH_BLOCKS) * len(T.BATCH_LANES) * len(T.BATCH_ROWS_PER_MEMBER)
    cfg = B.config()
    assert (cfg["block"], cfg["lanes_per_row"], cfg["rows_per_member"]) == \
        (best["block"], best["lanes_per_row"], best["rows_per_member"])
    best2, table2 = T.tune_batch(B, bd, max_its=500, rtol=1e-6, reps=1)
    assert table2 is None and best2 == best


def test_config_batch_full_size():
    """The full batch: 256 SPD systems, rows log-uniform in [1e3, 2e5] (9.9M
    rows), 4 right-hand sides each, one persistent launch to 1e-6; every
    right-hand side converges, and the true residuals of every 32nd system
    meet the tolerance."""
    nsys, nrhs = 256, 4
    ns = np.exp(np.log(1e3) + (np.log(2e5) - np.log(1e3)) * P.uniform(17, nsys)).astype(np.int64)
    systems = [P.random_spd(int(n), 1000 + i) for i, n in enumerate(ns)]
    B = S.Batch(systems, nrhs=nrhs)
    B.set_config(1024, 4, 4096)
    bh = rhs(B.nrows, nrhs, 23)
    x, its = B.solve(torch.from_numpy(bh).to(DEV), max_its=20000, rtol=1e-6)
    assert np.all(its > 0) and np.all(its < 20000)
    xh = x.cpu().numpy()
    for s in range(0, nsys, 32):
        A = systems[s]
        lo, hi = B.sys_row[s], B.sys_row[s + 1]
        oc = csr32(A)
        for j in range(nrhs):
            r = bh[lo:hi, j].astype(np.float64) - O.spmv(oc, xh[lo:hi, j].astype(np.float64))
            rel = np.linalg.norm(r) / np.linalg.norm(bh[lo:hi, j])
            assert meets_rtol(rel, oc, bh[lo:hi, j]), (s, j, rel)
