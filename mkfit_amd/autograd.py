"""torch.autograd view of the fit's total chi2 as a function of per-layer material and alignment.

    data = mk.AlignmentData(ctx, event, "cuda:0")
    theta = torch.zeros(n_layers, 3, device="cuda:0", requires_grad=True)   # (x/X0, d_rphi, d_z)
    loss = fit_chi2(theta, data)       # sum over tracks of the fit chi2 (mkf_fit_adjoint forward)
    loss.backward()                    # theta.grad from the same call's reverse sweep

The forward call already runs the reverse sweep (mkf_fit_adjoint returns both), so backward only
scales the stored gradient.
"""
import torch


class FitChi2(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, data):
        chi2, grad = data.run(theta)
        ctx.save_for_backward(grad.to(theta.dtype))
        return chi2.double().sum().to(theta.dtype)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None


def fit_chi2(theta, data):
    return FitChi2.apply(theta, data)
