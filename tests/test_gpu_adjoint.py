"""GPU checks of the adjoint (extension 3): theta -> A assembly, SDDMM
gradient, torch autograd against the CPU oracle's adjoint, central
differences and torch.autograd.gradcheck."""
import numpy as np
import pytest
import torch

import oracle as O
import spmv_amd as S
from spmv_amd import adjoint as AD
from spmv_amd import problems as P

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def problem(n, seed=1):
    ne = P.diffusion_num_edges(n)
    theta = np.exp(0.5 * P.normal(seed, ne))       # lognormal(0, 0.5)
    b = P.uniform(seed + 1, n * n, 0.0, 1.0)
    xt = P.normal(seed + 2, n * n)
    return theta, b, xt


def test_assembly_bitexact():
    n = 37
    theta, _, _ = problem(n)
    op = AD.DiffusionOperator(n)
    vals = op.assemble(torch.from_numpy(theta).to(DEV)).cpu().numpy()
    assert np.array_equal(vals, O.gen_diffusion(n, theta).val)


def test_gradients_match_oracle_adjoint():
    n = 48
    theta, b, xt = problem(n, 5)
    op = AD.DiffusionOperator(n)
    th = torch.from_numpy(theta).to(DEV).requires_grad_(True)
    bd = torch.from_numpy(b).to(DEV).requires_grad_(True)
    x = op.solve(th, bd, rtol=1e-14)
    loss = 0.5 * ((x - torch.from_numpy(xt).to(DEV)) ** 2).sum()
    loss.backward()
    ref = O.adjoint_diffusion(n, theta, b, xt, 100000, 1e-14)
    assert abs(loss.item() - ref["loss"]) <= 1e-10 * abs(ref["loss"])
    gt = th.grad.cpu().numpy()
    assert np.max(np.abs(gt - ref["grad_theta"])) <= 1e-9 * np.max(np.abs(ref["grad_theta"]))
    assert np.max(np.abs(bd.grad.cpu().numpy() - ref["grad_b"])) <= 1e-9 * np.max(np.abs(ref["grad_b"]))


def test_sddmm_gradient_is_minus_lambda_x():
    n = 20
    theta, b, xt = problem(n, 9)
    op = AD.DiffusionOperator(n)
    vals = op.assemble(torch.from_numpy(theta).to(DEV)).requires_grad_(True)
    x = AD.SparseSolve.apply(vals, torch.from_numpy(b).to(DEV), op.A, 100000, 1e-14)
    loss = 0.5 * ((x - torch.from_numpy(xt).to(DEV)) ** 2).sum()
    loss.backward()
    A = O.gen_diffusion(n, theta).to_scipy().toarray()
    xs = np.linalg.solve(A, b)
    lam = np.linalg.solve(A.T, xs - xt)
    pat = op.pattern
    rows = np.repeat(np.arange(pat.nrows), np.diff(pat.rowptr))
    want = -lam[rows] * xs[pat.colidx]
    assert np.allclose(vals.grad.cpu().numpy(), want, rtol=0, atol=1e-10 * np.abs(want).max())


def test_central_differences_32_entries():
    n = 64
    theta, b, xt = problem(n, 13)
    op = AD.DiffusionOperator(n)
    th = torch.from_numpy(theta).to(DEV).requires_grad_(True)
    bd = torch.from_numpy(b).to(DEV)
    xtd = torch.from_numpy(xt).to(DEV)
    x = op.solve(th, bd, rtol=1e-14)
    (0.5 * ((x - xtd) ** 2).sum()).backward()
    entries = (P.rand_u64(77, 32) % np.uint64(op.num_edges)).astype(np.int64)
    fd = AD.central_difference(op, th.detach(), bd, xtd, entries, eps=1e-6, rtol=1e-14)
    g = th.grad.cpu().numpy()[entries]
    assert np.max(np.abs(fd - g) / np.maximum(np.abs(g), 1e-300)) <= 1e-6


def test_gradcheck_tiny():
    n = 4
    theta, b, xt = problem(n, 21)
    op = AD.DiffusionOperator(n)
    xtd = torch.from_numpy(xt).to(DEV)

    def f(th, bb):
        x = op.solve(th, bb, rtol=1e-15)
        return 0.5 * ((x - xtd) ** 2).sum()

    th = torch.from_numpy(theta).to(DEV).requires_grad_(True)
    bd = torch.from_numpy(b).to(DEV).requires_grad_(True)
    assert torch.autograd.gradcheck(f, (th, bd), eps=1e-6, atol=1e-6, rtol=1e-5)


def test_config_diffusion_1024_full_size():
    """The 1024^2 anisotropic diffusion problem at full size (5,238,784
    entries): the adjoint solution satisfies A lambda = x - x_target to FP64
    CG's attainable accuracy, and
    central differences (eps 1e-6) on 4 seeded entries match the gradient to
    1e-6 relative."""
    n = 1024
    theta, b, xt = problem(n, 1)
    op = AD.DiffusionOperator(n)
    assert op.A.nnz == 5_238_784
    th = torch.from_numpy(theta).to(DEV).requires_grad_(True)
    bd = torch.from_numpy(b).to(DEV).requires_grad_(True)
    xtd = torch.from_numpy(xt).to(DEV)
    x = op.solve(th, bd, rtol=1e-13)
    (0.5 * ((x - xtd) ** 2).sum()).backward()
    lam = bd.grad.detach().cpu().numpy()          # grad_b = lambda
    A = O.gen_diffusion(n, theta)
    rhs_adj = (x.detach() - xtd).cpu().numpy()
    r = O.spmv(O.CSR(A.nrows, A.ncols, A.rowptr, A.colidx, A.val), lam, threads=8) - rhs_adj
    # the recursive residual stops at 1e-13; the true residual of FP64 CG on
    # this operator (condition number ~1e6-1e7) floors near kappa * eps ~ 1e-9
    assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(rhs_adj)
    entries = (P.rand_u64(77, 4) % np.uint64(op.num_edges)).astype(np.int64)
    fd = AD.central_difference(op, th.detach(), bd.detach(), xtd, entries, eps=1e-6, rtol=1e-13)
    g = th.grad.cpu().numpy()[entries]
    assert np.max(np.abs(fd - g) / np.maximum(np.abs(g), 1e-300)) <= 1e-6


def test_config_diffusion_1024_matches_oracle_adjoint():
    """The full-size 1024^2 problem against the oracle's adjoint on the host
    (the reference's CG for the forward and the adjoint solve, products and
    axpys over the host's threads, dot products in a fixed 64-chunk order):
    loss, grad_b = lambda and grad_theta = -(lambda_i - lambda_j)(x_i - x_j)
    per edge.  The two CG runs round their dot products differently (~8050 +
    ~7470 iterations on the GPU, 8065 + 7483 on the host), so they agree to
    1e-14 relative in x and lambda and 7e-12 in grad_theta, not bit for bit;
    the bounds hold a margin of ~1000."""
    import os
    n = 1024
    theta, b, xt = problem(n, 1)
    op = AD.DiffusionOperator(n)
    th = torch.from_numpy(theta).to(DEV).requires_grad_(True)
    bd = torch.from_numpy(b).to(DEV).requires_grad_(True)
    xtd = torch.from_numpy(xt).to(DEV)
    x = op.solve(th, bd, rtol=1e-13)
    loss = 0.5 * ((x - xtd) ** 2).sum()
    loss.backward()
    threads = max(1, min(16, os.cpu_count() or 1))
    A = O.gen_diffusion(n, theta)
    xo, k1, _ = O.cg(A, b, 100000, 1e-13, threads=threads)
    rhs = xo - xt
    lo = 0.5 * float(rhs @ rhs)
    lam, k2, _ = O.cg(A, rhs, 100000, 1e-13, threads=threads)
    rows = np.repeat(np.arange(A.nrows), np.diff(A.rowptr))
    go = O.diffusion_grad_theta(n, A, -lam[rows] * xo[A.colidx])
    xg = x.detach().cpu().numpy()
    lg = bd.grad.cpu().numpy()
    gt = th.grad.cpu().numpy()
    ex = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    el = np.linalg.norm(lg - lam) / np.linalg.norm(lam)
    eg = np.max(np.abs(gt - go)) / np.max(np.abs(go))
    print(f"iterations cpu {k1}+{k2}; rel diff x {ex:.2e} lambda {el:.2e} grad_theta {eg:.2e} "
          f"loss {abs(loss.item() - lo) / lo:.2e}", flush=True)
    assert abs(loss.item() - lo) <= 1e-12 * lo
    assert ex <= 1e-11 and el <= 1e-11 and eg <= 1e-8
