"""A thin torch.autograd.Function over the C-ABI (test wrapper): the
solvation energy of fixed charges' positions on a fixed surface as a
function of the charges and of kappa, with the backward through the
library's adjoint (fftsvd_problem_grad_charges, fftsvd_problem_grad_kappa).
CPU float64 tensors only: the library owns the GPU work.

The HIP library is loaded before torch, so torch's ROCm build and the
library share the one HIP runtime (same soname) that the library was
linked against."""
import fftsvd_amd as fa

fa.lib()
import torch  # noqa: E402


class Surface:
    """Panels, charge positions and the solve settings of one molecule."""

    def __init__(self, panels, qpos, eps_in=4.0, eps_out=80.0, tol=1e-10, precision=fa.F64, **problem_kw):
        self.panels, self.qpos = panels, qpos
        self.eps_in, self.eps_out = eps_in, eps_out
        self.tol, self.precision, self.kw = tol, precision, problem_kw


class SolvationEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, kappa, surf):
        prob = fa.Problem(surf.panels, surf.qpos, q.detach().numpy(), surf.eps_in, surf.eps_out,
                          float(kappa), **surf.kw)
        x, _, _, _ = prob.solve(prob.rhs(), surf.tol, precision=surf.precision)
        ctx.prob, ctx.x, ctx.surf = prob, x, surf
        return torch.tensor(prob.energy(x, 332.112), dtype=torch.float64)

    @staticmethod
    def backward(ctx, gE):
        prob, x, surf = ctx.prob, ctx.x, ctx.surf
        gq = torch.from_numpy(prob.grad_charges(x, surf.tol, precision=surf.precision))
        gk = prob.grad_kappa(x, surf.tol, precision=surf.precision)
        return gE * gq, gE * torch.tensor(gk, dtype=torch.float64), None


def solvation_energy(q, kappa, surf):
    return SolvationEnergy.apply(q, kappa, surf)
