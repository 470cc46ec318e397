"""torch.autograd.Function over the adjoint entry point (tests only): forward =
the aggregated multi-view NCC cost of each pixel's (theta, phi, depth), backward
= dpe_ncc_cost_adjoint's gradient.  Tensors live on cuda:0; the C-ABI gets
their device pointers."""
import ctypes as C

import torch

from dpe_amd import abi


class AggregatedNCC(torch.autograd.Function):
    @staticmethod
    def forward(ctx_, tpd, view_weights, dpe_ctx, params, inputs):
        """tpd: (H, W, 3) float32 cuda tensor of (theta, phi, depth); view_weights:
        (NV, H, W) float32 cuda tensor or None; inputs: abi.Inputs with
        memory = MEM_DEVICE.  Returns the (H, W) cost."""
        H, W, _ = tpd.shape
        planes = torch.zeros((H, W, 4), dtype=torch.float32, device=tpd.device)
        planes[..., :3] = tpd.detach()
        costs = torch.empty((H, W), dtype=torch.float32, device=tpd.device)
        grads = torch.empty((H, W, 4), dtype=torch.float32, device=tpd.device)
        vw = view_weights.contiguous() if view_weights is not None else None
        torch.cuda.synchronize()
        dpe_ctx.check(dpe_ctx.lib.dpe_ncc_cost_adjoint(
            dpe_ctx.h, C.byref(params), C.byref(inputs), planes.data_ptr(), vw.data_ptr() if vw is not None else None,
            costs.data_ptr(), grads.data_ptr()))
        ctx_.save_for_backward(grads[..., :3].contiguous())
        return costs

    @staticmethod
    def backward(ctx_, grad_cost):
        (g,) = ctx_.saved_tensors
        return grad_cost[..., None] * g, None, None, None, None


def device_inputs(pin, dev="cuda:0"):
    """abi.Inputs over cuda copies of a ProblemInputs' images (kept alive in the
    returned list)."""
    keep = [torch.from_numpy(im).to(dev) for im in pin.images]
    s = abi.Inputs()
    s.width, s.height, s.num_images, s.memory = pin.width, pin.height, len(pin.images), abi.MEM_DEVICE
    for i, t in enumerate(keep):
        s.images[i] = t.data_ptr()
        s.cameras[i] = pin.cameras[i]
    return s, keep
