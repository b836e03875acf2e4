"""Differentiation through the sparse solve (extension 3; no reference
counterpart).

For A(theta) x = b with A symmetric positive definite and a scalar loss
L(x), the adjoint method gives

    A^T lam = dL/dx            (same CG, A^T = A; transpose product of
                                spmv/Matrix.cpp:206-211 otherwise)
    dL/db   = lam
    dL/dA   = -lam x^T         restricted to the sparsity pattern (SDDMM)
    dL/dtheta = P^T vec(dL/dA) where A.values = P theta

Two torch.autograd.Functions, both running in the library's HIP kernels:

* :class:`SparseSolve`  values, b -> x   (CG forward, adjoint CG + SDDMM back)
* :class:`PatternAssemble` theta -> values = P theta (CSR SpMV forward, the
  transpose SpMV P^T g back)

:class:`DiffusionOperator` builds P for the 2D anisotropic 5-point diffusion
operator (edge conductivities theta, Dirichlet boundary edges) so that
``loss(solve(assemble(theta), b))`` is differentiable in theta and b.
"""
import numpy as np
import torch

from . import Matrix, SpmvError, cg, sddmm
from . import problems as P


class SparseSolve(torch.autograd.Function):
    """x = A(values)^{-1} b by CG on the device; backward by the adjoint."""

    @staticmethod
    def forward(ctx, values, b, A, max_its, rtol):
        if values.dtype != torch.float64 or b.dtype != torch.float64:
            raise SpmvError("SparseSolve works in float64")
        A.set_values(values.detach().contiguous())
        x, its = cg(A, b.detach().contiguous(), max_its, rtol)
        ctx.A, ctx.max_its, ctx.rtol = A, max_its, rtol
        ctx.its = [its]
        A.cg_iterations = {"forward": its}  # reporting only
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        A = ctx.A
        # A symmetric: A^T lam = grad_x is solved with the same CG.  A zero
        # seed has the exact adjoint lam = 0 (CG itself would form 0/0).
        if not torch.any(grad_x):
            lam = torch.zeros_like(x)
        else:
            lam, its = cg(A, grad_x.contiguous(), ctx.max_its, ctx.rtol)
            ctx.its.append(its)
            getattr(A, "cg_iterations", {})["adjoint"] = its
        grad_values = sddmm(A, lam, x, alpha=-1.0) if ctx.needs_input_grad[0] else None
        grad_b = lam if ctx.needs_input_grad[1] else None
        return grad_values, grad_b, None, None, None


class PatternAssemble(torch.autograd.Function):
    """values = P theta (P a CSR Matrix of coefficients); backward P^T g."""

    @staticmethod
    def forward(ctx, theta, Pm):
        ctx.Pm = Pm
        return Pm.mult(theta.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        return ctx.Pm.transpmult(g.contiguous()), None


class DiffusionOperator:
    """2D anisotropic 5-point diffusion on an n x n grid with edge
    conductivities theta (2n(n-1) interior + 4n Dirichlet boundary edges).

    A(theta) = sum_interior theta_e (e_a - e_b)(e_a - e_b)^T
             + sum_boundary theta_e e_a e_a^T
    """

    def __init__(self, n):
        self.n = n
        A, Pc = P.diffusion_pattern(n)
        self.pattern = A
        self.num_edges = P.diffusion_num_edges(n)
        self.A = Matrix(A.rowptr, A.colidx, np.zeros(A.nnz), ncols=A.ncols)
        self.Pm = Matrix(Pc.rowptr, Pc.colidx, Pc.val, ncols=Pc.ncols)

    def assemble(self, theta):
        return PatternAssemble.apply(theta, self.Pm)

    def solve(self, theta, b, max_its=100000, rtol=1e-12):
        values = self.assemble(theta)
        return SparseSolve.apply(values, b, self.A, max_its, rtol)


def central_difference(op, theta, b, x_target, entries, eps=1e-6, max_its=100000, rtol=1e-13):
    """dL/dtheta_e by central differences for L = 0.5||x - x_t||^2, for each
    e in `entries`.  L(theta+) - L(theta-) is evaluated without cancellation:
    with x- = x(theta - eps e_e) and d = x+ - x-, which solves
    A(theta+) d = -(theta+_e - theta-_e) P_e x-   (P_e = dA/dtheta_e),
    L+ - L- = d.(x- - x_t) + d.d / 2 exactly.  Returns a numpy array."""
    n = op.n
    nh = n * (n - 1)
    out = []
    with torch.no_grad():
        for e in entries:
            e = int(e)
            tp, tm = theta.clone(), theta.clone()
            tp[e] += eps
            tm[e] -= eps
            xm = op.solve(tm, b, max_its, rtol)
            # P_e x-: edge e joins nodes a-b (interior) or a alone (boundary)
            if e < nh:
                j, i = divmod(e, n - 1)
                a, c = i + n * j, i + 1 + n * j
            elif e < 2 * nh:
                j, i = divmod(e - nh, n)
                a, c = i + n * j, i + n * (j + 1)
            else:
                side, t = divmod(e - 2 * nh, n)
                a, c = [t, t + n * (n - 1), n * t, n - 1 + n * t][side], None
            dt = (tp[e] - tm[e]).item()
            rhs = torch.zeros_like(b)
            if c is None:
                rhs[a] = -dt * xm[a]
            else:
                diff = (xm[a] - xm[c]).item()
                rhs[a] = -dt * diff
                rhs[c] = dt * diff
            op.A.set_values(op.Pm.mult(tp))
            d, _ = cg(op.A, rhs, max_its, rtol)
            dL = torch.dot(d, xm - x_target) + 0.5 * torch.dot(d, d)
            out.append(dL.item() / dt)
    return np.array(out)
