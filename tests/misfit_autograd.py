"""Test-only torch.autograd.Function over the misfit forward/backward C-ABI
pair (hz_misfit_forward / hz_misfit_backward): J(m) = 1/2 sum_s ||P u_s - d_s||^2
with A(m) u_s = f_s, m = 1/c^2, and dJ/dm from the library's adjoint solves."""
import torch


class MisfitFunction(torch.autograd.Function):
    """J = MisfitFunction.apply(m, hz, desc, f, rec, data, prec, tol, store_fields).

    m: squared slowness over the grid (float64 tensor, any device); the
    problem is set up from c = 1/sqrt(m) and factored in `prec`."""

    @staticmethod
    def forward(ctx, m, hz, desc, f, rec, data, prec, tol, store_fields):
        c = (1.0 / torch.sqrt(m.detach())).double().cpu().numpy()
        S = hz.Solver(desc, c)
        S.factor(prec)
        J = S.misfit_forward(f, rec, data, m=40, restarts=50, tol=tol, prec=prec, store_fields=store_fields)
        ctx.solver = S
        ctx.m_device = m.device
        return torch.tensor(J, dtype=torch.float64, device=m.device)

    @staticmethod
    def backward(ctx, grad_J):
        g = ctx.solver.misfit_backward().to(ctx.m_device)
        ctx.solver.close()
        ctx.solver = None
        return (grad_J * g,) + (None,) * 8
