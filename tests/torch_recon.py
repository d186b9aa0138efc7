"""PyTorch-ROCm test harness around the C-ABI loss gradient: a torch.autograd.Function whose
forward evaluates loss = sum_b w_b P_b(delta_rec(rho)) and whose backward returns the hand-written
adjoint's dloss/drho (mmh_recon_power_grad_d).  Test infrastructure only."""
import torch


class ReconPowerLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, weights, lib_ctx, tol):
        g = torch.zeros_like(rho)
        loss, pk, st = lib_ctx.recon_power_grad(rho.detach().contiguous(), weights.detach().cpu().numpy(), g, tol=tol)
        ctx.save_for_backward(g)
        return rho.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None, None, None


def recon_power_loss(rho, weights, lib_ctx, tol=1e-12):
    return ReconPowerLoss.apply(rho, weights, lib_ctx, tol)
