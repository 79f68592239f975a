"""A thin torch.autograd.Function around the library's forward and adjoint,
used only as a test harness (the product is the C ABI): the forward gather of
a shot as a function of vp, whose backward is sfd_acoustic_gradient applied
to the incoming gradient of the gather (the seed of the vector-Jacobian
product).  CPU tensors; the library runs on the GPU."""
import numpy as np
import torch

import seismic_code_github_amd as sfd


class AcousticGather(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, shot, ckpt_every, sfd_ctx):
        v = vp.detach().cpu().numpy().astype(np.float32)
        s = sfd.AcousticShot(**{**shot.__dict__, "vp": v, "_keep": []})
        rec, _ = sfd.acoustic_forward(s, sfd_ctx)
        ctx.shot, ctx.ckpt_every, ctx.sfd_ctx = s, ckpt_every, sfd_ctx
        return torch.from_numpy(rec)

    @staticmethod
    def backward(ctx, grad_rec):
        seed = grad_rec.detach().cpu().numpy().astype(np.float32)
        g, _ = sfd.acoustic_gradient(ctx.shot, seed, ckpt_every=ctx.ckpt_every, ctx=ctx.sfd_ctx)
        return torch.from_numpy(g), None, None, None


def gather(vp, shot, ckpt_every=50, sfd_ctx=None):
    return AcousticGather.apply(vp, shot, ckpt_every, sfd_ctx or sfd.default_context())
