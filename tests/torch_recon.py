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


class ReconPowerLossMesh(torch.autograd.Function):
    """The loss at a curved mesh (mmh_recon_power_grad_mesh_d): inputs rho and def, both
    differentiable."""

    @staticmethod
    def forward(ctx, rho, defn, weights, lib_ctx, tol):
        gr, gd = torch.zeros_like(rho), torch.zeros_like(defn)
        loss, pk, st = lib_ctx.recon_power_grad_mesh(rho.detach().contiguous(), defn.detach().contiguous(),
                                                     weights.detach().cpu().numpy(), gr, gd, tol=tol)
        ctx.save_for_backward(gr, gd)
        return rho.new_tensor(loss)

    @staticmethod
    def backward(ctx, grad_out):
        gr, gd = ctx.saved_tensors
        return grad_out * gr, grad_out * gd, None, None, None


def recon_power_loss_mesh(rho, defn, weights, lib_ctx, tol=1e-12):
    return ReconPowerLossMesh.apply(rho, defn, weights, lib_ctx, tol)
