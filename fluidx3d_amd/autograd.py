"""torch.autograd.Function over the C ABI's forward/backward, used by the gradient-check
harness only: J = OutletFlux.apply(porosity, adjoint) with a float32 CUDA tensor of N
porosities; J.backward() fills porosity.grad with dJ/dporosity from fx_lbm_backward.

Import this module (it imports torch and starts torch's GPU runtime) before anything
loads libfx_lbm.so: torch and the library each ship libamdhip64.so.7, the first one
loaded serves the whole process, and torch's runtime does not come up on a copy it did
not load itself."""
import torch

if not torch.cuda.is_available():
    raise RuntimeError("torch sees no GPU; import fluidx3d_amd.autograd before the first fluidx3d_amd GPU call")
torch.cuda.init()


class OutletFlux(torch.autograd.Function):
    @staticmethod
    def forward(ctx, porosity, adjoint):
        lbm = adjoint.lbm
        if porosity.dtype != torch.float32 or not porosity.is_cuda or porosity.numel() != lbm.N:
            raise ValueError("porosity must be a float32 CUDA tensor with one value per cell")
        lbm.porosity.host[:] = porosity.detach().cpu().numpy()
        J = adjoint.forward()
        ctx.adjoint = adjoint
        ctx.device = porosity.device
        return torch.tensor(J, dtype=torch.float64, device=porosity.device)

    @staticmethod
    def backward(ctx, grad_J):
        adjoint = ctx.adjoint
        grad = torch.empty(adjoint.lbm.N, dtype=torch.float32, device=ctx.device)
        torch.cuda.synchronize(ctx.device)
        adjoint.backward_device(grad.data_ptr())  # synchronous on the context stream
        return (grad_J.to(torch.float32) * grad), None
