"""The adjoint behind torch.autograd (tests/torch_adjoint.py, a test harness):

* J = sum(seed * gather(vp)) differentiated by torch equals
  sfd_acoustic_gradient for that seed, bit for bit (the backward is that call);
* an L2 misfit over two shots written in torch gives the gradient
  sfd_acoustic_misfit_gradient computes natively (to 1e-6);
* the directional derivative of J from torch matches central finite
  differences of the FP64 oracle's J (to 1e-3).
"""
import numpy as np
import pytest
import torch

import oracle
import seismic_code_github_amd as sfd
from conftest import oracle_threads
from test_gpu_gradient import case
from torch_adjoint import gather

pytestmark = pytest.mark.gpu


def test_autograd_backward_is_the_c_abi_gradient(ctx):
    shot, _, sd = case(nt=180)
    print("case: autograd vs sfd_acoustic_gradient", flush=True)
    vp = torch.tensor(shot.vp, requires_grad=True)
    rec = gather(vp, shot, 40, ctx)
    J = (rec * torch.from_numpy(sd)).sum()
    J.backward()
    g, r = sfd.acoustic_gradient(shot, sd, ckpt_every=40, ctx=ctx)
    np.testing.assert_array_equal(rec.detach().numpy(), r)
    np.testing.assert_array_equal(vp.grad.numpy(), g)


def test_autograd_misfit_matches_native_misfit_gradient(ctx):
    print("case: autograd misfit over two shots", flush=True)
    shots, obs = [], []
    for sx in (25, 45):
        shot, _, _ = case(nt=160, src=(sx, 15))
        shots.append(shot)
        true = sfd.AcousticShot(**{**shot.__dict__, "vp": shot.vp * np.float32(1.02), "_keep": []})
        obs.append(sfd.acoustic_forward(true, ctx)[0])
    vp = torch.tensor(shots[0].vp, requires_grad=True)
    loss = sum(0.5 * ((gather(vp, s, 40, ctx) - torch.from_numpy(d)) ** 2).sum() for s, d in zip(shots, obs))
    loss.backward()
    J, g = sfd.misfit_gradient(shots, obs, ckpt_every=40, ctx=ctx)
    assert abs(float(loss.detach()) - J) <= 1e-5 * J
    e = float(np.linalg.norm(vp.grad.numpy().astype(np.float64) - g) / np.linalg.norm(g))
    print(f"torch misfit gradient vs native: rel-L2 {e:.2e}")
    assert e <= 1e-6


def test_autograd_directional_derivative_vs_fp64_finite_differences(ctx):
    from scipy.ndimage import gaussian_filter
    shot, okw, sd = case(nt=200)
    print("case: autograd directional derivative", flush=True)
    vp = torch.tensor(shot.vp, requires_grad=True)
    (gather(vp, shot, 50, ctx) * torch.from_numpy(sd)).sum().backward()
    d = gaussian_filter(np.random.default_rng(5).standard_normal(shot.vp.shape), 3)

    def J(v):
        o = oracle.acoustic(**{**okw, "vp": v}, precision="f64", nthreads=oracle_threads())
        return float(np.sum(sd.astype(np.float64) * o["record"]))

    v0 = shot.vp.astype(np.float64)
    vpp = (v0 + 1e-2 * d).astype(np.float32)
    vpm = (v0 - 1e-2 * d).astype(np.float32)
    step = (vpp.astype(np.float64) - vpm.astype(np.float64)) / 2
    fd = (J(vpp) - J(vpm)) / 2
    an = float(np.sum(vp.grad.numpy().astype(np.float64) * step))
    print(f"fd {fd:.8e} autograd {an:.8e} rel {abs(fd - an) / abs(fd):.2e}")
    assert abs(fd - an) <= 1e-3 * abs(fd)
