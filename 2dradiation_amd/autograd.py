"""torch.autograd wrapper of the GPU stream-collide adjoint.

``ObserverRing.apply(eta, kappa0, ctx, I0, nSteps, every, x, y)`` runs nSteps
fused steps from I0 with the per-cell emissivity eta and absorption kappa0
(float64 CPU tensors of the stored-grid size; the other fields stay as last
set on ctx through ``ctx.set_fields`` - this function replaces eta and kappa0
only) and returns the outward intensities sampled at the points (x, y)
[npts, nDir].  Its backward calls the library's checkpointed adjoint
(include/rt2d.h: rt_forward_checkpointed / rt_adjoint), so gradients reach
eta and kappa0 without torch seeing the intermediate states.  The context
holds the checkpoints of one forward at a time: backward of an output whose
forward has since been replaced by another one on the same context raises.
"""
import numpy as np
import torch


class ObserverRing(torch.autograd.Function):
    @staticmethod
    def forward(actx, eta, kappa0, ctx, I0, nSteps, every, x, y, fields):
        f = dict(fields)
        f["eta"] = eta.detach().cpu().double().numpy()
        f["kappa0"] = kappa0.detach().cpu().double().numpy()
        ctx.set_fields(**f)
        ctx.upload_I(I0)
        ctx.forward_checkpointed(nSteps, every)
        out = ctx.sample_outgoing(x, y)
        actx.rt = (ctx, np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64),
                   ctx.checkpoint_generation)
        return torch.from_numpy(out)

    @staticmethod
    def backward(actx, grad_out):
        ctx, x, y, gen = actx.rt
        if ctx.checkpoint_generation != gen:
            raise RuntimeError("ObserverRing.backward: the context ran another checkpointed forward since this "
                               "output was computed; its checkpoints belong to that run")
        ge, gk = ctx.adjoint(x, y, grad_out.detach().cpu().double().numpy())
        return torch.from_numpy(ge), torch.from_numpy(gk), None, None, None, None, None, None, None


def observer_ring(eta, kappa0, ctx, I0, nSteps, every, x, y, fields=None):
    """Differentiable observer-ring samples; see ObserverRing."""
    return ObserverRing.apply(eta, kappa0, ctx, I0, nSteps, every, x, y, fields or {})
