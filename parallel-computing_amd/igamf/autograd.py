"""torch.autograd bindings of the kappa-weighted operator and solve.

``apply_kappa(op, x, kappa)``  y = A(kappa) x
    backward: grad_x = A^T g (iga_op_apply_backward), grad_kappa[e][q] =
    -w_q det(J_q) grad x_q . grad (P_int g)_q (hand-written HIP kernel).
``solve_kappa(solver, f, kappa)``  u = A(kappa)^-1 f (interior unknowns)
    backward (adjoint): lambda = A^-T g, grad_f = lambda,
    grad_kappa = -(d/dkappa lambda^T A(kappa) u) via the same kernel.

Both run on the HIP library only (no CPU path); kappa is [E][nqp] with the
reference's quadrature order, FP64.
"""
import torch


class _ApplyKappa(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, kappa, op):
        op.set_kappa(kappa.detach())
        y = op.apply(x.detach().contiguous())
        ctx.save_for_backward(x, kappa)
        ctx.op = op
        return y

    @staticmethod
    def backward(ctx, gy):
        x, kappa = ctx.saved_tensors
        op = ctx.op
        op.set_kappa(kappa.detach())
        want_x, want_k = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gx, gk = op.backward(x.detach().contiguous(), gy.contiguous(), want_x=want_x,
                             want_kappa=want_k)
        if gk is not None:
            gk = gk.view_as(kappa)
        return gx, gk, None


def apply_kappa(op, x, kappa):
    return _ApplyKappa.apply(x, kappa, op)


class _SolveKappa(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, kappa, solver):
        op = solver.op
        op.set_kappa(kappa.detach())
        u, st = solver.solve(f.detach().contiguous())
        ctx.save_for_backward(u, kappa)
        ctx.solver = solver
        ctx.stats = st
        return u

    @staticmethod
    def backward(ctx, gu):
        u, kappa = ctx.saved_tensors
        solver = ctx.solver
        op = solver.op
        op.set_kappa(kappa.detach())
        # the interior block of A is symmetric: A^T lambda = g on interior rows
        g = gu.contiguous() * interior_mask(op)
        lam, _ = solver.solve(g)
        gk = None
        if ctx.needs_input_grad[1]:
            _, gk = op.backward(u, lam, want_x=False, want_kappa=True)
            gk = -gk.view_as(kappa)
        return lam, gk, None


def interior_mask(op):
    if getattr(op, "_interior", None) is None:
        idv = torch.from_numpy(op.mesh.arrays()["ID"]).to(op.device)
        op._interior = (idv >= 0).to(torch.float64)
    return op._interior


def solve_kappa(solver, f, kappa):
    return _SolveKappa.apply(f, kappa, solver)
